// ORACLE (test infrastructure only): constraint rows and the Newton solver of one environment.
// Restates genesis/engine/solvers/rigid/constraint/solver.py (the batch/CPU code path, "monolith" solve body).
#pragma once
#include "oracle_dynamics.h"
#include "oracle_model.h"

namespace orc {

// jacobian row of a point p attached to `link`, component i of sign * (cdof_vel - (p - root_COM) x cdof_ang), added
// into row c along the link's chain (dofs of each link last to first); returns the row's contribution to J v.
// with_rot: the velocity goes through qd_transform_motion_by_trans_quat with the identity rotation (connect rows,
// solver.py:757-760), else it is taken as is (weld rows, :1086-1089).
template <class T>
T point_jac_row(const Model<T>& m, Env<T>& s, int c, int link, V3<T> p, T sign, int i, bool with_rot) {
    using E = Env<T>;
    const int D = m.D;
    T jq = 0;
    while (link > -1) {
        for (int dd = 0; dd < m.link_dof_end[link] - m.link_dof_start[link]; dd++) {
            int k = m.link_dof_end[link] - 1 - dd;
            V3<T> ca = E::g3(s.cdof_ang, k), cv = E::g3(s.cdof_vel, k);
            V3<T> tpos = p - E::g3(s.root_COM, link);
            V3<T> vel = cv - cross(tpos, ca);
            if (with_rot) vel = transform_by_quat(vel, identity_quat<T>());
            V3<T> diff = sign * vel;
            T j = diff[i];
            jq = jq + j * s.vel[k];
            s.jac[(size_t)c * D + k] = s.jac[(size_t)c * D + k] + j;
        }
        link = m.link_parent[link];
    }
    return jq;
}

// solver.py:689-797 func_equality_connect: three rows holding the two anchors together
template <class T> void equality_connect(const Model<T>& m, Env<T>& s, int i_e) {
    using E = Env<T>;
    const int D = m.D;
    const T* sp = &m.eq_sol_params[7 * i_e];
    const T* data = &m.eq_data[11 * i_e];
    const int l1 = m.eq_obj1[i_e], l2 = m.eq_obj2[i_e];
    V3<T> a1{data[0], data[1], data[2]}, a2{data[3], data[4], data[5]};
    V3<T> g1 = transform_by_trans_quat(a1, E::g3(s.links_pos, l1), E::g4(s.links_quat, l1));
    V3<T> g2 = transform_by_trans_quat(a2, E::g3(s.links_pos, l2), E::g4(s.links_quat, l2));
    T invw = m.link_invweight[2 * l1] + m.link_invweight[2 * l2];
    for (int i3 = 0; i3 < 3; i3++) {
        int c = s.n_con++;
        s.n_con_eq++;
        for (int d = 0; d < D; d++) s.jac[(size_t)c * D + d] = T(0);
        T jq = 0;
        for (int i_ab = 0; i_ab < 2; i_ab++) {
            T sign = i_ab == 0 ? T(1) : T(-1);
            jq = jq + point_jac_row(m, s, c, i_ab == 0 ? l1 : l2, i_ab == 0 ? g1 : g2, sign, i3, true);
        }
        V3<T> pd = g1 - g2;
        T pen = norm(pd);
        T imp, aref;
        imp_aref(sp, -pen, jq, pd[i3], imp, aref);
        T dg = std::fmax(invw * (T(1) - imp) / imp, m.EPS);
        s.diag[c] = dg;
        s.aref[c] = aref;
        s.efc_D[c] = T(1) / dg;
    }
}

// solver.py:984-1173 func_equality_weld: three position rows, then three orientation rows
template <class T> void equality_weld(const Model<T>& m, Env<T>& s, int i_e) {
    using E = Env<T>;
    const int D = m.D;
    const T* sp = &m.eq_sol_params[7 * i_e];
    const T* data = &m.eq_data[11 * i_e];
    const int l1 = m.eq_obj1[i_e], l2 = m.eq_obj2[i_e];
    V3<T> a1{data[3], data[4], data[5]}, a2{data[0], data[1], data[2]};
    Q4<T> relpose{data[6], data[7], data[8], data[9]};
    const T ts = data[10];
    Q4<T> q1 = E::g4(s.links_quat, l1), q2 = E::g4(s.links_quat, l2);
    V3<T> g1 = transform_by_trans_quat(a1, E::g3(s.links_pos, l1), q1);
    V3<T> g2 = transform_by_trans_quat(a2, E::g3(s.links_pos, l2), q2);
    V3<T> pe = g1 - g2;
    Q4<T> q = quat_mul(q1, relpose);
    Q4<T> iq2 = inv_quat(q2);
    Q4<T> eq = quat_mul(iq2, q);
    V3<T> re = V3<T>{eq.x, eq.y, eq.z} * ts;
    T all[6] = {pe.x, pe.y, pe.z, re.x, re.y, re.z};
    T sq = 0;
    for (int k = 0; k < 6; k++) sq = sq + all[k] * all[k];
    const T pos_imp = std::sqrt(sq);
    const T invw0 = m.link_invweight[2 * l1] + m.link_invweight[2 * l2];
    const T invw1 = m.link_invweight[2 * l1 + 1] + m.link_invweight[2 * l2 + 1];
    for (int i = 0; i < 3; i++) {
        int c = s.n_con++;
        s.n_con_eq++;
        for (int d = 0; d < D; d++) s.jac[(size_t)c * D + d] = T(0);
        T jq = 0;
        for (int i_ab = 0; i_ab < 2; i_ab++) {
            T sign = i_ab == 0 ? T(1) : T(-1);
            jq = jq + point_jac_row(m, s, c, i_ab == 0 ? l1 : l2, i_ab == 0 ? g1 : g2, sign, i, false);
        }
        T imp, aref;
        imp_aref(sp, -pos_imp, jq, pe[i], imp, aref);
        T dg = std::fmax(invw0 * (T(1) - imp) / imp, m.EPS);
        s.diag[c] = dg;
        s.aref[c] = aref;
        s.efc_D[c] = T(1) / dg;
    }
    const int c0 = s.n_con;
    s.n_con += 3;
    s.n_con_eq += 3;
    for (int k = 0; k < 3; k++)
        for (int d = 0; d < D; d++) s.jac[(size_t)(c0 + k) * D + d] = T(0);
    for (int i_ab = 0; i_ab < 2; i_ab++) {
        T sign = i_ab == 0 ? T(1) : T(-1);
        int link = i_ab == 0 ? l1 : l2;
        while (link > -1) {
            for (int dd = 0; dd < m.link_dof_end[link] - m.link_dof_start[link]; dd++) {
                int d = m.link_dof_end[link] - 1 - dd;
                V3<T> jv = sign * E::g3(s.cdof_ang, d);
                for (int k = 0; k < 3; k++) s.jac[(size_t)(c0 + k) * D + d] = s.jac[(size_t)(c0 + k) * D + d] + jv[k];
            }
            link = m.link_parent[link];
        }
    }
    T jq3[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
        V3<T> jd{s.jac[(size_t)c0 * D + d], s.jac[(size_t)(c0 + 1) * D + d], s.jac[(size_t)(c0 + 2) * D + d]};
        Q4<T> quat2 = quat_mul_axis(iq2, jd);
        Q4<T> quat3 = quat_mul(quat2, q);
        for (int k = 0; k < 3; k++) {
            s.jac[(size_t)(c0 + k) * D + d] = T(0.5) * quat3[k + 1] * ts;
            jq3[k] = jq3[k] + s.jac[(size_t)(c0 + k) * D + d] * s.vel[d];
        }
    }
    for (int k = 0; k < 3; k++) {
        T imp, aref;
        imp_aref(sp, -pos_imp, jq3[k], re[k], imp, aref);
        T dg = std::fmax(invw1 * (T(1) - imp) / imp, m.EPS);
        s.diag[c0 + k] = dg;
        s.aref[c0 + k] = aref;
        s.efc_D[c0 + k] = T(1) / dg;
    }
}

// solver.py:801-878 func_equality_joint
template <class T> void equality_joint(const Model<T>& m, Env<T>& s, int i_e) {
    const int D = m.D;
    const T* sp = &m.eq_sol_params[7 * i_e];
    const T* data = &m.eq_data[11 * i_e];
    int j1 = m.eq_obj1[i_e], j2 = m.eq_obj2[i_e];
    int q1 = m.joint_q_start[j1], q2 = m.joint_q_start[j2];
    int d1 = m.joint_dof_start[j1], d2 = m.joint_dof_start[j2];
    int c = s.n_con++;
    s.n_con_eq++;
    for (int d = 0; d < D; d++) s.jac[(size_t)c * D + d] = T(0);
    T diff = s.qpos[q2] - m.qpos0[q2];
    T pos = s.qpos[q1] - m.qpos0[q1];
    T deriv = 0;
    for (int i5 = 0; i5 < 5; i5++) {
        T dp = std::pow(diff, T(i5));
        pos = pos - dp * data[i5];
        if (i5 < 4) deriv = deriv + data[i5 + 1] * dp * T(i5 + 1);
    }
    s.jac[(size_t)c * D + d1] = T(1);
    s.jac[(size_t)c * D + d2] = -deriv;
    T jac_qvel = s.jac[(size_t)c * D + d1] * s.vel[d1] + s.jac[(size_t)c * D + d2] * s.vel[d2];
    T invw = m.dof_invweight[d1] + m.dof_invweight[d2];
    T imp, aref;
    imp_aref(sp, -std::fabs(pos), jac_qvel, pos, imp, aref);
    T dg = std::fmax(invw * (T(1) - imp) / imp, m.EPS);
    s.diag[c] = dg;
    s.aref[c] = aref;
    s.efc_D[c] = T(1) / dg;
}

// solver.py:881-938 add_equality_constraints: the rows of every equality, in equality order
template <class T> void add_equality_constraints(const Model<T>& m, Env<T>& s) {
    s.n_con = 0;
    s.n_con_eq = 0;
    for (int i_e = 0; i_e < m.NE; i_e++) {
        if (m.eq_type[i_e] == EQ_CONNECT) equality_connect(m, s, i_e);
        else if (m.eq_type[i_e] == EQ_WELD) equality_weld(m, s, i_e);
        else if (m.eq_type[i_e] == EQ_JOINT) equality_joint(m, s, i_e);
    }
}

// solver.py:1237-1286 add_frictionloss_constraints
template <class T> void add_frictionloss_constraints(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    s.n_con_fl = 0;
    for (int i_l = 0; i_l < m.L; i_l++)
        for (int i_j = m.link_joint_start[i_l]; i_j < m.link_joint_end[i_l]; i_j++)
            for (int d = m.joint_dof_start[i_j]; d < m.joint_dof_end[i_j]; d++) {
                if (!(m.dof_frictionloss[d] > m.EPS)) continue;
                T jac_qvel = T(1) * s.vel[d];
                T imp, aref;
                imp_aref(&m.joint_sol_params[7 * i_j], T(0), jac_qvel, T(0), imp, aref);
                T dg = std::fmax(m.dof_invweight[d] * (T(1) - imp) / imp, m.EPS);
                int c = s.n_con++;
                s.n_con_fl++;
                s.diag[c] = dg;
                s.aref[c] = aref;
                s.efc_D[c] = T(1) / dg;
                s.efc_frictionloss[c] = m.dof_frictionloss[d];
                for (int k = 0; k < D; k++) s.jac[(size_t)c * D + k] = T(0);
                s.jac[(size_t)c * D + d] = T(1);
            }
}

// solver.py:580-685 add_collision_constraints (4 pyramid rows per contact)
template <class T> void add_collision_constraints(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    const int D = m.D;
    int start = s.n_con;
    for (int i_col = 0; i_col < s.n_contacts; i_col++) {
        int la = s.contact_link_a[i_col], lb = s.contact_link_b[i_col];
        V3<T> cpos = E::g3(s.contact_pos, i_col);
        V3<T> nrm = E::g3(s.contact_normal, i_col);
        T fr = s.contact_friction[i_col];
        const T* sp = &s.contact_sol_params[7 * i_col];
        T pen = s.contact_penetration[i_col];
        V3<T> d1, d2;
        orthogonals(nrm, d1, d2);
        T invw = m.link_invweight[2 * la];
        if (lb > -1) invw = invw + m.link_invweight[2 * lb];
        for (int i = 0; i < 4; i++) {
            V3<T> d = T(2 * (i % 2) - 1) * (i < 2 ? d1 : d2);
            V3<T> n = d * fr - nrm;
            int c = start + i_col * 4 + i;
            for (int k = 0; k < D; k++) s.jac[(size_t)c * D + k] = T(0);
            T jac_qvel = 0;
            for (int i_ab = 0; i_ab < 2; i_ab++) {
                T sign = i_ab == 0 ? T(-1) : T(1);
                int link = i_ab == 0 ? la : lb;
                while (link > -1) {
                    for (int dd = 0; dd < m.link_dof_end[link] - m.link_dof_start[link]; dd++) {
                        int k = m.link_dof_end[link] - 1 - dd;
                        V3<T> ca = E::g3(s.cdof_ang, k), cv = E::g3(s.cdof_vel, k);
                        V3<T> tpos = cpos - E::g3(s.root_COM, link);
                        // qd_transform_motion_by_trans_quat with identity rotation
                        V3<T> vel = transform_by_quat(cv - cross(tpos, ca), identity_quat<T>());
                        V3<T> diff = sign * vel;
                        T j = dot(diff, n);
                        jac_qvel = jac_qvel + j * s.vel[k];
                        s.jac[(size_t)c * D + k] = s.jac[(size_t)c * D + k] + j;
                    }
                    link = m.link_parent[link];
                }
            }
            T imp, aref;
            imp_aref(sp, -pen, jac_qvel, -pen, imp, aref);
            T dg = invw + fr * fr * invw;
            dg *= T(2) * fr * fr * (T(1) - imp) / imp;
            dg = std::fmax(dg, m.EPS);
            s.diag[c] = dg;
            s.aref[c] = aref;
            s.efc_D[c] = T(1) / dg;
        }
    }
    s.n_con = s.n_con + s.n_contacts * 4;
}

// solver.py:1177-1233 add_joint_limit_constraints
template <class T> void add_joint_limit_constraints(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    for (int i_l = 0; i_l < m.L; i_l++)
        for (int i_j = m.link_joint_start[i_l]; i_j < m.link_joint_end[i_l]; i_j++) {
            int jt = m.joint_type[i_j];
            if (jt != JT_REVOLUTE && jt != JT_PRISMATIC) continue;
            int q = m.joint_q_start[i_j], d = m.joint_dof_start[i_j];
            T dmin = s.qpos[q] - m.dof_limit[2 * d];
            T dmax = m.dof_limit[2 * d + 1] - s.qpos[q];
            T pd = std::fmin(dmin, dmax);
            if (pd < 0) {
                T j = T((dmin < dmax) ? 1 : 0) * 2 - 1;
                T jac_qvel = j * s.vel[d];
                T imp, aref;
                imp_aref(&m.joint_sol_params[7 * i_j], pd, jac_qvel, pd, imp, aref);
                T dg = std::fmax(m.dof_invweight[d] * (T(1) - imp) / imp, m.EPS);
                int c = s.n_con++;
                s.diag[c] = dg;
                s.aref[c] = aref;
                s.efc_D[c] = T(1) / dg;
                for (int k = 0; k < D; k++) s.jac[(size_t)c * D + k] = T(0);
                s.jac[(size_t)c * D + d] = j;
            }
        }
}

template <class T> void add_inequality_constraints(const Model<T>& m, Env<T>& s) {
    add_frictionloss_constraints(m, s);
    if (m.h.enable_collision) add_collision_constraints(m, s);
    if (m.h.enable_joint_limit) add_joint_limit_constraints(m, s);
}

// solver.py:2620-2694 func_update_constraint_batch
template <class T>
void update_constraint(const Model<T>& m, Env<T>& s, const std::vector<T>& qacc, const std::vector<T>& Ma, T& cost) {
    const int D = m.D;
    int ne = s.n_con_eq, nef = ne + s.n_con_fl;
    s.prev_cost = cost;
    T cost_i = 0, gauss_i = 0;
    for (int c = 0; c < s.n_con; c++) {
        s.prev_active[c] = s.active[c];
        s.active[c] = 1;
        T floss_force = 0;
        if (ne <= c && c < nef) {
            T f = s.efc_frictionloss[c], r = s.diag[c];
            T rf = r * f;
            bool ln = s.Jaref[c] <= -rf, lp = s.Jaref[c] >= rf;
            s.active[c] = !(ln || lp);
            floss_force = T(ln) * f + T(lp) * -f;
            T fc = T(ln) * f * (T(-0.5) * rf - s.Jaref[c]);
            fc = fc + T(lp) * f * (T(-0.5) * rf + s.Jaref[c]);
            cost_i = cost_i + fc;
        } else if (nef <= c) {
            s.active[c] = s.Jaref[c] < 0;
        }
        s.efc_force[c] = floss_force + (-s.Jaref[c] * s.efc_D[c] * T(s.active[c]));
    }
    for (int d = 0; d < D; d++) {
        T q = 0;
        for (int c = 0; c < s.n_con; c++) q = q + s.jac[(size_t)c * D + d] * s.efc_force[c];
        s.qfrc_constraint[d] = q;
    }
    for (int d = 0; d < D; d++) {
        T v = T(0.5) * (Ma[d] - s.force[d]) * (qacc[d] - s.acc_smooth[d]);
        gauss_i = gauss_i + v;
        cost_i = cost_i + v;
    }
    for (int c = 0; c < s.n_con; c++)
        cost_i = cost_i + T(0.5) * (s.Jaref[c] * s.Jaref[c] * s.efc_D[c] * T(s.active[c]));
    s.gauss = gauss_i;
    cost = cost_i;
}

// solver.py:1395-1452 (dense) + :1595-1621
template <class T> void hessian_cholesky_direct(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    for (int i = 0; i < D; i++)
        for (int j = 0; j <= i; j++) s.nt_H[i * D + j] = T(0);
    for (int i = 0; i < D; i++)
        for (int c = 0; c < s.n_con; c++) {
            T ji = s.jac[(size_t)c * D + i];
            if (std::fabs(ji) > m.EPS)
                for (int j = 0; j <= i; j++)
                    s.nt_H[i * D + j] = s.nt_H[i * D + j] + s.jac[(size_t)c * D + j] * ji * s.efc_D[c] * T(s.active[c]);
        }
    for (int i_e = 0; i_e < m.E; i_e++)
        for (int i = m.entity_dof_start[i_e]; i < m.entity_dof_end[i_e]; i++)
            for (int j = m.entity_dof_start[i_e]; j <= i; j++) s.nt_H[i * D + j] = s.nt_H[i * D + j] + s.mass_mat[i * D + j];
    for (int i = 0; i < D; i++) {
        T tmp = s.nt_H[i * D + i];
        for (int j = 0; j < i; j++) tmp = tmp - s.nt_H[i * D + j] * s.nt_H[i * D + j];
        s.nt_H[i * D + i] = std::sqrt(std::fmax(tmp, m.EPS));
        tmp = T(1) / s.nt_H[i * D + i];
        for (int j = i + 1; j < D; j++) {
            T dt = 0;
            for (int k = 0; k < i; k++) dt = dt + s.nt_H[j * D + k] * s.nt_H[i * D + k];
            s.nt_H[j * D + i] = (s.nt_H[j * D + i] - dt) * tmp;
        }
    }
}

// solver.py:1764-1822 incremental rank-1 update/downdate; returns true if degenerate
template <class T> bool hessian_cholesky_incremental(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    for (int c = 0; c < s.n_con; c++) {
        if (!(s.active[c] ^ s.prev_active[c])) continue;
        T sign = s.active[c] ? T(1) : T(-1);
        T sq = std::sqrt(s.efc_D[c]);
        for (int d = 0; d < D; d++) s.nt_vec[d] = s.jac[(size_t)c * D + d] * sq;
        for (int k = 0; k < D; k++) {
            if (!(std::fabs(s.nt_vec[k]) > m.EPS)) continue;
            T Lkk = s.nt_H[k * D + k];
            T tmp = Lkk * Lkk + sign * s.nt_vec[k] * s.nt_vec[k];
            if (tmp < m.EPS) return true;
            T r = std::sqrt(tmp);
            T cc = r / Lkk;
            T cinv = T(1) / cc;
            T ss = s.nt_vec[k] / Lkk;
            s.nt_H[k * D + k] = r;
            for (int i = k + 1; i < D; i++) s.nt_H[i * D + k] = (s.nt_H[i * D + k] + ss * s.nt_vec[i] * sign) * cinv;
            for (int i = k + 1; i < D; i++) s.nt_vec[i] = s.nt_vec[i] * cc - ss * s.nt_H[i * D + k];
        }
    }
    return false;
}

// solver.py:1893-1910 func_cholesky_solve_batch; :2722-2751 gradient
template <class T> void update_gradient(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    for (int d = 0; d < D; d++) s.grad[d] = s.Ma[d] - s.force[d] - s.qfrc_constraint[d];
    if (m.h.solver_type == SOLVER_CG) {  // :2737-2747 Mgrad = M^-1 grad (func_solve_mass_batch)
        solve_mass(m, s, s.grad, s.Mgrad);
        return;
    }
    for (int i = 0; i < D; i++) {
        T cur = s.grad[i];
        for (int j = 0; j < i; j++) cur = cur - s.nt_H[i * D + j] * s.Mgrad[j];
        s.Mgrad[i] = cur / s.nt_H[i * D + i];
    }
    for (int ii = 0; ii < D; ii++) {
        int i = D - 1 - ii;
        T cur = s.Mgrad[i];
        for (int j = i + 1; j < D; j++) cur = cur - s.nt_H[j * D + i] * s.Mgrad[j];
        s.Mgrad[i] = cur / s.nt_H[i * D + i];
    }
}

template <class T> struct LsPoint {
    T alpha, cost, grad, hess;
};

// solver.py:2034-2152 func_ls_init_and_eval_p0_opt
template <class T> LsPoint<T> ls_init_p0(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    int ne = s.n_con_eq, nef = ne + s.n_con_fl, nc = s.n_con;
    for (int i_e = 0; i_e < m.E; i_e++)
        for (int i = m.entity_dof_start[i_e]; i < m.entity_dof_end[i_e]; i++) {
            T mv = 0;
            for (int j = m.entity_dof_start[i_e]; j < m.entity_dof_end[i_e]; j++) mv = mv + s.mass_mat[i * D + j] * s.search[j];
            s.mv[i] = mv;
        }
    for (int c = 0; c < nc; c++) {
        T jv = 0;
        for (int d = 0; d < D; d++) jv = jv + s.jac[(size_t)c * D + d] * s.search[d];
        s.jv[c] = jv;
    }
    T qg1 = 0, qg2 = 0;
    for (int d = 0; d < D; d++) {
        qg1 = qg1 + (s.search[d] * s.Ma[d] - s.search[d] * s.force[d]);
        qg2 = qg2 + T(0.5) * s.search[d] * s.mv[d];
    }
    s.quad_gauss[0] = s.gauss;
    s.quad_gauss[1] = qg1;
    s.quad_gauss[2] = qg2;
    T t0 = s.gauss, t1 = qg1, t2 = qg2;
    T e0 = 0, e1 = 0, e2 = 0;
    for (int c = 0; c < nc; c++) {
        T Ja = s.Jaref[c], jv = s.jv[c], Dc = s.efc_D[c];
        T f0 = Dc * (T(0.5) * Ja * Ja), f1 = Dc * (jv * Ja), f2 = Dc * (T(0.5) * jv * jv);
        if (c < ne) {
            e0 = e0 + f0; e1 = e1 + f1; e2 = e2 + f2;
            t0 = t0 + f0; t1 = t1 + f1; t2 = t2 + f2;
        } else if (c < nef) {
            T f = s.efc_frictionloss[c], r = s.diag[c];
            T rf = r * f;
            bool ln = Ja <= -rf, lp = Ja >= rf;
            if (ln || lp) {
                f0 = T(ln) * f * (T(-0.5) * rf - Ja) + T(lp) * f * (T(-0.5) * rf + Ja);
                f1 = T(ln) * (-f * jv) + T(lp) * (f * jv);
                f2 = 0;
            }
            t0 = t0 + f0; t1 = t1 + f1; t2 = t2 + f2;
        } else {
            T a = T(Ja < 0);
            t0 = t0 + f0 * a; t1 = t1 + f1 * a; t2 = t2 + f2 * a;
        }
    }
    s.eq_sum[0] = e0; s.eq_sum[1] = e1; s.eq_sum[2] = e2;
    T hess = T(2) * t2;
    if (hess <= T(0)) hess = m.EPS;
    s.ls_it = 1;
    return {T(0), t0, t1, hess};
}

// solver.py:2155-2223 func_ls_point_fn_opt (evaluates n alphas in one pass; n = 1 or 3)
template <class T> void ls_eval(const Model<T>& m, Env<T>& s, const T* alphas, int n, LsPoint<T>* out) {
    int ne = s.n_con_eq, nef = ne + s.n_con_fl, nc = s.n_con;
    T b0 = s.quad_gauss[0] + s.eq_sum[0], b1 = s.quad_gauss[1] + s.eq_sum[1], b2 = s.quad_gauss[2] + s.eq_sum[2];
    T t[3][3];
    for (int k = 0; k < n; k++) { t[k][0] = b0; t[k][1] = b1; t[k][2] = b2; }
    for (int c = ne; c < nef; c++) {
        T Ja = s.Jaref[c], jv = s.jv[c], Dc = s.efc_D[c], f = s.efc_frictionloss[c], r = s.diag[c];
        T f0 = Dc * (T(0.5) * Ja * Ja), f1 = Dc * (jv * Ja), f2 = Dc * (T(0.5) * jv * jv);
        T rf = r * f;
        for (int k = 0; k < n; k++) {
            T x = Ja + alphas[k] * jv;
            bool ln = x <= -rf, lp = x >= rf;
            T a0 = f0, a1 = f1, a2 = f2;
            if (ln || lp) {
                a0 = T(ln) * f * (T(-0.5) * rf - Ja) + T(lp) * f * (T(-0.5) * rf + Ja);
                a1 = T(ln) * (-f * jv) + T(lp) * (f * jv);
                a2 = 0;
            }
            t[k][0] = t[k][0] + a0; t[k][1] = t[k][1] + a1; t[k][2] = t[k][2] + a2;
        }
    }
    for (int c = nef; c < nc; c++) {
        T Ja = s.Jaref[c], jv = s.jv[c], Dc = s.efc_D[c];
        T f0 = Dc * (T(0.5) * Ja * Ja), f1 = Dc * (jv * Ja), f2 = Dc * (T(0.5) * jv * jv);
        for (int k = 0; k < n; k++) {
            T x = Ja + alphas[k] * jv;
            T a = T(x < 0);
            t[k][0] = t[k][0] + f0 * a; t[k][1] = t[k][1] + f1 * a; t[k][2] = t[k][2] + f2 * a;
        }
    }
    for (int k = 0; k < n; k++) {
        T al = alphas[k];
        T cost = al * al * t[k][2] + al * t[k][1] + t[k][0];
        T grad = T(2) * al * t[k][2] + t[k][1];
        T hess = T(2) * t[k][2];
        if (hess <= T(0)) hess = m.EPS;
        out[k] = {al, cost, grad, hess};
    }
    s.ls_it += n;
}

// solver.py:2358-2389
template <class T> int update_bracket(LsPoint<T>& p, const LsPoint<T>* c, T& next_alpha) {
    int flag = 0;
    for (int i = 0; i < 3; i++) {
        if (p.grad < 0 && c[i].grad < 0 && p.grad < c[i].grad) {
            p = c[i];
            flag = 1;
        } else if (p.grad > 0 && c[i].grad > 0 && p.grad > c[i].grad) {
            p = c[i];
            flag = 2;
        }
    }
    next_alpha = p.alpha;
    if (flag > 0) next_alpha = p.alpha - p.grad / p.hess;
    return flag;
}

// solver.py:2427-2597 func_linesearch_batch
template <class T> T linesearch(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    T snorm = 0;
    for (int d = 0; d < D; d++) snorm = snorm + s.search[d] * s.search[d];
    snorm = std::sqrt(snorm);
    T scale = T(m.meaninertia) * T(std::max(1, D));
    T gtol = T(m.tolerance) * T(m.ls_tolerance) * snorm * scale;
    s.gtol = gtol;
    s.ls_it = 0;
    s.ls_result = 0;
    const int ls_iters = m.h.ls_iterations;
    T res = 0;
    if (snorm < m.EPS) {
        s.ls_result = 1;
        return T(0);
    }
    LsPoint<T> p0 = ls_init_p0(m, s);
    LsPoint<T> p1;
    T a = p0.alpha - p0.grad / p0.hess;
    ls_eval(m, s, &a, 1, &p1);
    if (p0.cost < p1.cost) p1 = p0;
    if (std::fabs(p1.grad) < gtol) {
        s.ls_result = std::fabs(p1.alpha) < m.EPS ? 2 : 0;
        return p1.alpha;
    }
    T direction = T((p1.grad < 0) ? 1 : 0) * 2 - 1;
    int p2update = 0;
    LsPoint<T> p2 = p1;
    while (p1.grad * direction <= -gtol && s.ls_it < ls_iters) {
        p2 = p1;
        p2update = 1;
        T an = p1.alpha - p1.grad / p1.hess;
        ls_eval(m, s, &an, 1, &p1);
        if (std::fabs(p1.grad) < gtol) return p1.alpha;
    }
    if (s.ls_it >= ls_iters) {
        s.ls_result = 3;
        return p1.alpha;
    }
    if (!p2update) {
        s.ls_result = 6;
        return p1.alpha;
    }
    T al[3] = {p1.alpha - p1.grad / p1.hess, p1.alpha, (p1.alpha + p2.alpha) * T(0.5)};
    while (s.ls_it < ls_iters) {
        LsPoint<T> c[3];
        ls_eval(m, s, al, 3, c);
        T n1 = al[0], n2 = al[1];
        bool found = false;
        T best_a = 0, best_c = 0;
        for (int i = 0; i < 3; i++)
            if (std::fabs(c[i].grad) < gtol && (!found || c[i].cost < best_c)) {
                best_a = c[i].alpha;
                best_c = c[i].cost;
                found = true;
            }
        if (found) return best_a;
        int b1 = update_bracket(p1, c, n1);
        int b2 = update_bracket(p2, c, n2);
        if (b1 == 0 && b2 == 0) {
            s.ls_result = c[2].cost < p0.cost ? 0 : 7;
            return al[2];
        }
        al[0] = n1;
        al[1] = n2;
        al[2] = (p1.alpha + p2.alpha) * T(0.5);
    }
    if (p1.cost <= p2.cost && p1.cost < p0.cost) {
        s.ls_result = 4;
        res = p1.alpha;
    } else if (p2.cost <= p1.cost && p2.cost < p0.cost) {
        s.ls_result = 4;
        res = p2.alpha;
    } else {
        s.ls_result = 5;
        res = T(0);
    }
    return res;
}

// solver.py:2936-3059 func_solve_init, :3063-3139 func_solve_iter (Newton or CG), :3164-3190 monolith body,
// :3241-3261 func_update_qacc, :3199-3237 func_update_contact_force
template <class T> void constraint_solve(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    for (int d = 0; d < D; d++) s.qacc[d] = (s.n_con > 0 && s.is_warmstart) ? s.qacc_ws[d] : s.acc_smooth[d];
    for (int d1 = 0; d1 < D; d1++) {
        int e = m.dof_entity[d1];
        T ma = 0;
        for (int d2 = m.entity_dof_start[e]; d2 < m.entity_dof_end[e]; d2++) ma = ma + s.mass_mat[d1 * D + d2] * s.qacc[d2];
        s.Ma[d1] = ma;
    }
    for (int c = 0; c < s.n_con; c++) {
        T ja = -s.aref[c];
        for (int d = 0; d < D; d++) ja = ja + s.jac[(size_t)c * D + d] * s.qacc[d];
        s.Jaref[c] = ja;
    }
    update_constraint(m, s, s.qacc, s.Ma, s.cost);
    s.improved = s.n_con > 0;
    const bool cg = m.h.solver_type == SOLVER_CG;
    if (!cg) hessian_cholesky_direct(m, s);
    update_gradient(m, s);
    for (int d = 0; d < D; d++) s.search[d] = -s.Mgrad[d];
    s.n_iter = 0;
    if (s.n_con > 0) {
        for (int it = 0; it < m.h.iterations; it++) {
            s.n_iter++;
            T alpha = linesearch(m, s);
            if (std::fabs(alpha) < m.EPS) {
                s.improved = 0;
            } else {
                for (int d = 0; d < D; d++) {
                    s.qacc[d] = s.qacc[d] + s.search[d] * alpha;
                    s.Ma[d] = s.Ma[d] + s.mv[d] * alpha;
                }
                for (int c = 0; c < s.n_con; c++) s.Jaref[c] = s.Jaref[c] + s.jv[c] * alpha;
                if (cg)
                    for (int d = 0; d < D; d++) {
                        s.cg_prev_grad[d] = s.grad[d];
                        s.cg_prev_Mgrad[d] = s.Mgrad[d];
                    }
                update_constraint(m, s, s.qacc, s.Ma, s.cost);
                if (!cg && hessian_cholesky_incremental(m, s)) hessian_cholesky_direct(m, s);
                update_gradient(m, s);
                T tol = (T(m.meaninertia) * T(std::max(1, D))) * T(m.tolerance);
                T improvement = s.prev_cost - s.cost;
                T gn = 0;
                for (int d = 0; d < D; d++) gn = gn + s.grad[d] * s.grad[d];
                gn = std::sqrt(gn);
                s.improved = gn > tol && improvement > tol;
                if (s.improved && !cg) {
                    for (int d = 0; d < D; d++) s.search[d] = -s.Mgrad[d];
                } else if (s.improved) {  // :2865-2884 Polak-Ribiere direction
                    T beta = 0, pg_dot_pMg = 0;
                    for (int d = 0; d < D; d++) {
                        beta = beta + s.grad[d] * (s.Mgrad[d] - s.cg_prev_Mgrad[d]);
                        pg_dot_pMg = pg_dot_pMg + s.cg_prev_Mgrad[d] * s.cg_prev_grad[d];
                    }
                    beta = std::fmax(beta / std::fmax(m.EPS, pg_dot_pMg), T(0));
                    s.cg_pg_dot_pMg = pg_dot_pMg;
                    s.cg_beta = beta;
                    for (int d = 0; d < D; d++) s.search[d] = -s.Mgrad[d] + beta * s.search[d];
                }
            }
            if (!s.improved) break;
        }
    } else {
        s.improved = 0;
    }
    for (int d = 0; d < D; d++) {
        s.acc[d] = s.qacc[d];
        s.qf_constraint[d] = s.qfrc_constraint[d];
        s.force[d] = s.qf_smooth[d] + s.qfrc_constraint[d];
        s.qacc_ws[d] = s.qacc[d];
        if (std::isnan(s.qacc[d])) s.errno_ |= RS_ERR_INVALID_FORCE_NAN;
    }
    s.is_warmstart = 1;
    int cs = s.n_con_eq + s.n_con_fl;
    for (int i = 0; i < s.n_contacts; i++) {
        V3<T> nrm = Env<T>::g3(s.contact_normal, i);
        T fr = s.contact_friction[i];
        V3<T> d1, d2, f = v3(T(0), T(0), T(0));
        orthogonals(nrm, d1, d2);
        for (int k = 0; k < 4; k++) {
            V3<T> d = T(2 * (k % 2) - 1) * (k < 2 ? d1 : d2);
            V3<T> n = d * fr - nrm;
            f = f + n * s.efc_force[i * 4 + k + cs];
        }
        Env<T>::s3(s.contact_force, i, f);
    }
}

}  // namespace orc
