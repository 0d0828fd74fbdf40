// Constraint rows and Newton solver of one environment (one lane per environment, dense rows in [row][env] planes).
// Restates genesis/engine/solvers/rigid/constraint/solver.py on its batch code path ("monolith" solve body,
// solver.py:3164-3190) with the exact piecewise-quadratic line search (solver.py:2427-2597).
#pragma once
#include "rs_device.h"
#include "rs_dyn.h"

namespace rs {

#define JAC(c, d) FP(st, jac)[((size_t)(c) * m.D + (d)) * st.B + b]

// Row store of the per-stage kernels: the [row][env] planes of the slab.
struct PlaneRows {
    const DState& st;
    int b, D;
    __device__ __forceinline__ float& jac(int c, int d) const { return FP(st, jac)[((size_t)c * D + d) * st.B + b]; }
    __device__ __forceinline__ float& diag(int c) const { return FP(st, diag)[(size_t)c * st.B + b]; }
    __device__ __forceinline__ float& aref(int c) const { return FP(st, aref)[(size_t)c * st.B + b]; }
    __device__ __forceinline__ float& efc_D(int c) const { return FP(st, efc_D)[(size_t)c * st.B + b]; }
    __device__ __forceinline__ float& fl(int c) const { return FP(st, efc_frictionloss)[(size_t)c * st.B + b]; }
};

// Rows are built from the kinematic state (st, b); contact records are read from (cs, cb).

// jacobian row of a point p attached to `link`, component i of sign * (cdof_vel - (p - root_COM) x cdof_ang), added into
// row c along the link's chain (dofs of each link last to first); returns the row's contribution to J v. kRot: the
// velocity goes through the identity-rotation transform of solver.py:757-760 (connect rows), else it is taken as is
// (weld rows, :1086-1089).
template <bool kRot, class R, class S>
__device__ inline float point_jac_row(const DModel& m, const S& st, int b, const R& rows, int c, int link, f3 p,
                                      float sign, int i) {
    float jq = 0.f;
    while (link > -1) {
        for (int k = m.link_dof_end[link] - 1; k >= m.link_dof_start[link]; k--) {
            f3 ca = LD3(cdof_ang, k), cv = LD3(cdof_vel, k);
            f3 tpos = p - LD3(root_COM, link);
            f3 vel = cv - cross3(tpos, ca);
            if (kRot) vel = qrot(vel, qident());
            const float j = (sign * vel)[i];
            jq = jq + j * SF(vel, k);
            rows.jac(c, k) = rows.jac(c, k) + j;
        }
        link = m.link_parent[link];
    }
    return jq;
}

// solver.py:689-797 func_equality_connect: three rows holding the two anchors together
template <class R, class S>
__device__ inline void connect_rows(const DModel& m, const S& st, int b, const R& rows, int e, int& nc) {
    const int D = m.D;
    cfptr sp = &m.eq_sol_params[7 * e];
    cfptr data = &m.eq_data[11 * e];
    const int l1 = m.eq_obj1[e], l2 = m.eq_obj2[e];
    const f3 g1 = tr_point(mk3(data[0], data[1], data[2]), LD3(links_pos, l1), LD4(links_quat, l1));
    const f3 g2 = tr_point(mk3(data[3], data[4], data[5]), LD3(links_pos, l2), LD4(links_quat, l2));
    const float invw = m.link_invweight[2 * l1] + m.link_invweight[2 * l2];
    for (int i3 = 0; i3 < 3; i3++) {
        const int c = nc++;
        for (int d = 0; d < D; d++) rows.jac(c, d) = 0.f;
        float jq = 0.f;
        jq = jq + point_jac_row<true>(m, st, b, rows, c, l1, g1, 1.f, i3);
        jq = jq + point_jac_row<true>(m, st, b, rows, c, l2, g2, -1.f, i3);
        const f3 pd = g1 - g2;
        const float pen = norm3(pd);
        float imp, aref;
        imp_aref(sp, -pen, jq, pd[i3], imp, aref);
        const float dg = fmaxf(invw * (1.f - imp) / imp, m.eps);
        rows.diag(c) = dg;
        rows.aref(c) = aref;
        rows.efc_D(c) = 1.f / dg;
    }
}

// solver.py:984-1173 func_equality_weld: three position rows, then three orientation rows
template <class R, class S>
__device__ inline void weld_rows(const DModel& m, const S& st, int b, const R& rows, int e, int& nc) {
    const int D = m.D;
    cfptr sp = &m.eq_sol_params[7 * e];
    cfptr data = &m.eq_data[11 * e];
    const int l1 = m.eq_obj1[e], l2 = m.eq_obj2[e];
    const f4q relpose{data[6], data[7], data[8], data[9]};
    const float ts = data[10];
    const f4q q1 = LD4(links_quat, l1), q2 = LD4(links_quat, l2);
    const f3 g1 = tr_point(mk3(data[3], data[4], data[5]), LD3(links_pos, l1), q1);
    const f3 g2 = tr_point(mk3(data[0], data[1], data[2]), LD3(links_pos, l2), q2);
    const f3 pe = g1 - g2;
    const f4q q = qmul(q1, relpose);
    const f4q iq2 = qinv(q2);
    const f4q eq = qmul(iq2, q);
    const f3 re = mk3(eq.x, eq.y, eq.z) * ts;
    const float all[6] = {pe.x, pe.y, pe.z, re.x, re.y, re.z};
    float sq = 0.f;
    for (int k = 0; k < 6; k++) sq = sq + all[k] * all[k];
    const float pos_imp = sqrtf(sq);
    const float invw0 = m.link_invweight[2 * l1] + m.link_invweight[2 * l2];
    const float invw1 = m.link_invweight[2 * l1 + 1] + m.link_invweight[2 * l2 + 1];
    for (int i = 0; i < 3; i++) {
        const int c = nc++;
        for (int d = 0; d < D; d++) rows.jac(c, d) = 0.f;
        float jq = 0.f;
        jq = jq + point_jac_row<false>(m, st, b, rows, c, l1, g1, 1.f, i);
        jq = jq + point_jac_row<false>(m, st, b, rows, c, l2, g2, -1.f, i);
        float imp, aref;
        imp_aref(sp, -pos_imp, jq, pe[i], imp, aref);
        const float dg = fmaxf(invw0 * (1.f - imp) / imp, m.eps);
        rows.diag(c) = dg;
        rows.aref(c) = aref;
        rows.efc_D(c) = 1.f / dg;
    }
    const int c0 = nc;
    nc += 3;
    for (int k = 0; k < 3; k++)
        for (int d = 0; d < D; d++) rows.jac(c0 + k, d) = 0.f;
    for (int ab = 0; ab < 2; ab++) {
        const float sign = ab == 0 ? 1.f : -1.f;
        int link = ab == 0 ? l1 : l2;
        while (link > -1) {
            for (int d = m.link_dof_end[link] - 1; d >= m.link_dof_start[link]; d--) {
                const f3 jv = sign * LD3(cdof_ang, d);
                for (int k = 0; k < 3; k++) rows.jac(c0 + k, d) = rows.jac(c0 + k, d) + jv[k];
            }
            link = m.link_parent[link];
        }
    }
    float jq3[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < D; d++) {
        const f3 jd = mk3(rows.jac(c0, d), rows.jac(c0 + 1, d), rows.jac(c0 + 2, d));
        const f4q quat3 = qmul(qmul_axis(iq2, jd), q);
        const float qv[3] = {quat3.x, quat3.y, quat3.z};
        for (int k = 0; k < 3; k++) {
            rows.jac(c0 + k, d) = 0.5f * qv[k] * ts;
            jq3[k] = jq3[k] + rows.jac(c0 + k, d) * SF(vel, d);
        }
    }
    for (int k = 0; k < 3; k++) {
        float imp, aref;
        imp_aref(sp, -pos_imp, jq3[k], re[k], imp, aref);
        const float dg = fmaxf(invw1 * (1.f - imp) / imp, m.eps);
        rows.diag(c0 + k) = dg;
        rows.aref(c0 + k) = aref;
        rows.efc_D(c0 + k) = 1.f / dg;
    }
}

// solver.py:801-878 func_equality_joint
template <class R, class S>
__device__ inline void joint_equality_row(const DModel& m, const S& st, int b, const R& rows, int e, int& nc) {
    const int D = m.D;
    cfptr sp = &m.eq_sol_params[7 * e];
    cfptr data = &m.eq_data[11 * e];
    int j1 = m.eq_obj1[e], j2 = m.eq_obj2[e];
    int q1 = m.joint_q_start[j1], q2 = m.joint_q_start[j2];
    int d1 = m.joint_dof_start[j1], d2 = m.joint_dof_start[j2];
    int c = nc++;
    for (int d = 0; d < D; d++) rows.jac(c, d) = 0.f;
    float diff = SF(qpos, q2) - m.qpos0[q2];
    float pos = SF(qpos, q1) - m.qpos0[q1];
    float deriv = 0.f, dp = 1.f;
    for (int i5 = 0; i5 < 5; i5++) {
        pos = pos - dp * data[i5];
        if (i5 < 4) deriv = deriv + data[i5 + 1] * dp * (float)(i5 + 1);
        dp = dp * diff;
    }
    rows.jac(c, d1) = 1.f;
    rows.jac(c, d2) = -deriv;
    float jq = rows.jac(c, d1) * SF(vel, d1) + rows.jac(c, d2) * SF(vel, d2);
    float invw = m.dof_invweight[d1] + m.dof_invweight[d2];
    float imp, aref;
    imp_aref(sp, -fabsf(pos), jq, pos, imp, aref);
    float dg = fmaxf(invw * (1.f - imp) / imp, m.eps);
    rows.diag(c) = dg;
    rows.aref(c) = aref;
    rows.efc_D(c) = 1.f / dg;
}

// solver.py:881-938 add_equality_constraints: the rows of every equality in equality order (connect 3, weld 6,
// joint 1); returns the number of rows
template <class R, class S>
__device__ inline int equality_rows(const DModel& m, const S& st, int b, const R& rows) {
    int nc = 0;
    for (int e = 0; e < m.NE; e++) {
        const int t = m.eq_type[e];
        if (t == EQ_CONNECT) connect_rows(m, st, b, rows, e, nc);
        else if (t == EQ_WELD) weld_rows(m, st, b, rows, e, nc);
        else if (t == EQ_JOINT) joint_equality_row(m, st, b, rows, e, nc);
    }
    return nc;
}

// solver.py:1237-1286 add_frictionloss_constraints; appends from row nc, returns the number of rows added
template <class R, class S>
__device__ inline int frictionloss_rows(const DModel& m, const S& st, int b, const R& rows, int nc) {
    const int D = m.D;
    int nfl = 0;
    for (int l = 0; l < m.L; l++)
        for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++)
            for (int d = m.joint_dof_start[j]; d < m.joint_dof_end[j]; d++) {
                if (!(m.dof_frictionloss[d] > m.eps)) continue;
                float imp, aref;
                imp_aref(&m.joint_sol_params[7 * j], 0.f, 1.f * SF(vel, d), 0.f, imp, aref);
                float dg = fmaxf(m.dof_invweight[d] * (1.f - imp) / imp, m.eps);
                int c = nc + nfl++;
                rows.diag(c) = dg;
                rows.aref(c) = aref;
                rows.efc_D(c) = 1.f / dg;
                rows.fl(c) = m.dof_frictionloss[d];
                for (int k = 0; k < D; k++) rows.jac(c, k) = 0.f;
                rows.jac(c, d) = 1.f;
            }
    return nfl;
}

// Where contact_rows accumulates a jacobian row: in place in the row store (RowJac), or in a staging column that is
// copied to the row store once the row is complete (the fused kernel stages in LDS, so that the walk's
// read-modify-writes do not go through global memory).
struct RowJac {
    template <class R>
    __device__ __forceinline__ float& at(const R& rows, int c, int k) const {
        return rows.jac(c, k);
    }
    template <class R>
    __device__ __forceinline__ void flush(const R&, int, int) const {}
};

// The kinematic tree as the jacobian walk reads it (parent link, dof range of a link): from the model (ChainModel), or
// from a copy that the fused kernel stages in LDS (ChainLds: parent, dof start, dof end, L entries each), since the
// walk's reads at per-lane addresses are a chain of dependent loads.
struct ChainModel {
    const DModel& m;
    __device__ __forceinline__ int parent(int l) const { return m.link_parent[l]; }
    __device__ __forceinline__ int dof_start(int l) const { return m.link_dof_start[l]; }
    __device__ __forceinline__ int dof_end(int l) const { return m.link_dof_end[l]; }
};
struct ChainLds {
    const int* p;
    int L;
    __device__ __forceinline__ int parent(int l) const { return p[l]; }
    __device__ __forceinline__ int dof_start(int l) const { return p[L + l]; }
    __device__ __forceinline__ int dof_end(int l) const { return p[2 * L + l]; }
};

// solver.py:580-685 add_collision_constraints, row i (0..3) of contact ic: row start + 4 ic + i (pyramid direction
// +-d1, +-d2 from the contact normal). The impedance depends on the contact only.
template <class R, class S, class J, class C>
__device__ inline void contact_row(const DModel& m, const S& st, int b, const DState& cs, int cb, const R& rows,
                                   int ic, int i, int start, const J& jacc, const C& chain) {
    const int D = m.D;
    const size_t CB = (size_t)cs.B;
    int la = cs.contact_link_a[ic * CB + cb], lb = cs.contact_link_b[ic * CB + cb];
    f3 cpos = ld3(cs.contact_pos, ic, cs.B, cb), nrm = ld3(cs.contact_normal, ic, cs.B, cb);
    float fr = cs.contact_friction[ic * CB + cb], pen = cs.contact_penetration[ic * CB + cb];
    float sp[7];
    for (int k = 0; k < 7; k++) sp[k] = cs.contact_sol_params[(size_t)(7 * ic + k) * CB + cb];
    f3 d1, d2;
    orthogonals(nrm, d1, d2);
    float invw = m.link_invweight[2 * la];
    if (lb > -1) invw = invw + m.link_invweight[2 * lb];
    const float imp = impedance(sp, -pen);
    f3 dd = (float)(2 * (i % 2) - 1) * (i < 2 ? d1 : d2);
    f3 n = dd * fr - nrm;
    int c = start + ic * 4 + i;
    for (int k = 0; k < D; k++) jacc.at(rows, c, k) = 0.f;
    float jq = 0.f;
    for (int ab = 0; ab < 2; ab++) {
        float sign = ab == 0 ? -1.f : 1.f;
        int link = ab == 0 ? la : lb;
        while (link > -1) {
            f3 tpos = cpos - LD3(root_COM, link);
            const int k0 = chain.dof_start(link), k1 = chain.dof_end(link), lp = chain.parent(link);
            for (int k = k1 - 1; k >= k0; k--) {
                f3 ca = LD3(cdof_ang, k), cv = LD3(cdof_vel, k);
                f3 vel = qrot(cv - cross3(tpos, ca), qident());
                float jj = dot3(sign * vel, n);
                jq = jq + jj * SF(vel, k);
                jacc.at(rows, c, k) = jacc.at(rows, c, k) + jj;
            }
            link = lp;
        }
    }
    jacc.flush(rows, c, D);
    const float aref = reference_acc(sp, imp, jq, -pen);
    float dg = invw + fr * fr * invw;
    dg *= 2.f * fr * fr * (1.f - imp) / imp;
    dg = fmaxf(dg, m.eps);
    rows.diag(c) = dg;
    rows.aref(c) = aref;
    rows.efc_D(c) = 1.f / dg;
}

template <class R, class S, class J = RowJac>
__device__ inline void contact_row(const DModel& m, const S& st, int b, const DState& cs, int cb, const R& rows,
                                   int ic, int i, int start, const J& jacc = J()) {
    contact_row(m, st, b, cs, cb, rows, ic, i, start, jacc, ChainModel{m});
}

// the four rows of contact ic: rows start + 4 ic .. + 3
template <class R, class S, class J = RowJac>
__device__ inline void contact_rows(const DModel& m, const S& st, int b, const DState& cs, int cb, const R& rows,
                                    int ic, int start, const J& jacc = J()) {
    for (int i = 0; i < 4; i++) contact_row(m, st, b, cs, cb, rows, ic, i, start, jacc);
}

// solver.py:1177-1233 add_joint_limit_constraints; appends from row nc, returns the new row count
template <class R, class S>
__device__ inline int limit_rows(const DModel& m, const S& st, int b, const R& rows, int nc) {
    const int D = m.D;
    for (int l = 0; l < m.L; l++)
        for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
            int jt = m.joint_type[j];
            if (jt != JT_REVOLUTE && jt != JT_PRISMATIC) continue;
            int q = m.joint_q_start[j], d = m.joint_dof_start[j];
            float dmin = SF(qpos, q) - m.dof_limit[2 * d];
            float dmax = m.dof_limit[2 * d + 1] - SF(qpos, q);
            float pd = fminf(dmin, dmax);
            if (pd < 0.f) {
                float jj = (float)((dmin < dmax) ? 1 : 0) * 2.f - 1.f;
                float imp, aref;
                imp_aref(&m.joint_sol_params[7 * j], pd, jj * SF(vel, d), pd, imp, aref);
                float dg = fmaxf(m.dof_invweight[d] * (1.f - imp) / imp, m.eps);
                int c = nc++;
                rows.diag(c) = dg;
                rows.aref(c) = aref;
                rows.efc_D(c) = 1.f / dg;
                for (int k = 0; k < D; k++) rows.jac(c, k) = 0.f;
                rows.jac(c, d) = jj;
            }
        }
    return nc;
}

template <class S>
__device__ inline void dev_equality_constraints(const DModel& m, const S& st, int b) {
    int nc = equality_rows(m, st, b, PlaneRows{st, b, m.D});
    SF(n_con, 0) = nc;
    SF(n_con_eq, 0) = nc;
}

template <class S>
__device__ inline void dev_inequality_constraints(const DModel& m, const S& st, int b) {
    PlaneRows rows{st, b, m.D};
    int nc = SF(n_con, 0);
    int nfl = frictionloss_rows(m, st, b, rows, nc);
    nc += nfl;
    SF(n_con_fl, 0) = nfl;
    if (m.enable_collision) {
        int ncont = SF(n_contacts, 0);
        for (int ic = 0; ic < ncont; ic++) contact_rows(m, st, b, st, b, rows, ic, nc);
        nc = nc + ncont * 4;
    }
    if (m.enable_joint_limit) nc = limit_rows(m, st, b, rows, nc);
    SF(n_con, 0) = nc;
}

// solver.py:2620-2694
template <class S>
__device__ inline float dev_update_constraint(const DModel& m, const S& st, int b, float cost_in) {
    const int D = m.D;
    int nc = SF(n_con, 0), ne = SF(n_con_eq, 0), nef = ne + SF(n_con_fl, 0);
    SF(prev_cost, 0) = cost_in;
    float cost = 0.f, gauss = 0.f;
    for (int c = 0; c < nc; c++) {
        SF(prev_active, c) = SF(active, c);
        int act = 1;
        float ff = 0.f;
        float ja = SF(Jaref, c);
        if (ne <= c && c < nef) {
            float f = SF(efc_frictionloss, c), r = SF(diag, c);
            float rf = r * f;
            bool ln = ja <= -rf, lp = ja >= rf;
            act = !(ln || lp);
            ff = (float)ln * f + (float)lp * -f;
            float fc = (float)ln * f * (-0.5f * rf - ja);
            fc = fc + (float)lp * f * (-0.5f * rf + ja);
            cost = cost + fc;
        } else if (nef <= c) {
            act = ja < 0.f;
        }
        SF(active, c) = act;
        SF(efc_force, c) = ff + (-ja * SF(efc_D, c) * (float)act);
    }
    for (int d = 0; d < D; d++) {
        float q = 0.f;
        for (int c = 0; c < nc; c++) q = q + JAC(c, d) * SF(efc_force, c);
        SF(qfrc_constraint, d) = q;
    }
    for (int d = 0; d < D; d++) {
        float v = 0.5f * (SF(Ma, d) - SF(force, d)) * (SF(qacc, d) - SF(acc_smooth, d));
        gauss = gauss + v;
        cost = cost + v;
    }
    for (int c = 0; c < nc; c++) {
        float ja = SF(Jaref, c);
        cost = cost + 0.5f * (ja * ja * SF(efc_D, c) * (float)SF(active, c));
    }
    SF(gauss, 0) = gauss;
    return cost;
}

// solver.py:1395-1452 + :1595-1621
template <class S>
__device__ inline void dev_hessian_cholesky(const DModel& m, const S& st, int b) {
    const int D = m.D;
    int nc = SF(n_con, 0);
    for (int i = 0; i < D; i++)
        for (int j = 0; j <= i; j++) SF(nt_H, i * D + j) = 0.f;
    for (int i = 0; i < D; i++)
        for (int c = 0; c < nc; c++) {
            float ji = JAC(c, i);
            if (fabsf(ji) > m.eps) {
                float w = SF(efc_D, c) * (float)SF(active, c);
                for (int j = 0; j <= i; j++) SF(nt_H, i * D + j) = SF(nt_H, i * D + j) + JAC(c, j) * ji * SF(efc_D, c) * (float)SF(active, c);
                (void)w;
            }
        }
    for (int e = 0; e < m.E; e++)
        for (int i = m.entity_dof_start[e]; i < m.entity_dof_end[e]; i++)
            for (int j = m.entity_dof_start[e]; j <= i; j++) SF(nt_H, i * D + j) = SF(nt_H, i * D + j) + SF(mass_mat, i * D + j);
    for (int i = 0; i < D; i++) {
        float t = SF(nt_H, i * D + i);
        for (int j = 0; j < i; j++) t = t - SF(nt_H, i * D + j) * SF(nt_H, i * D + j);
        float dii = sqrtf(fmaxf(t, m.eps));
        SF(nt_H, i * D + i) = dii;
        float inv = 1.f / dii;
        for (int j = i + 1; j < D; j++) {
            float dt = 0.f;
            for (int k = 0; k < i; k++) dt = dt + SF(nt_H, j * D + k) * SF(nt_H, i * D + k);
            SF(nt_H, j * D + i) = (SF(nt_H, j * D + i) - dt) * inv;
        }
    }
}

// solver.py:1782-1822; returns true when the update degenerates
template <class S>
__device__ inline bool dev_cholesky_incremental(const DModel& m, const S& st, int b) {
    const int D = m.D;
    int nc = SF(n_con, 0);
    for (int c = 0; c < nc; c++) {
        if (!(SF(active, c) ^ SF(prev_active, c))) continue;
        float sign = SF(active, c) ? 1.f : -1.f;
        float sq = sqrtf(SF(efc_D, c));
        for (int d = 0; d < D; d++) SF(nt_vec, d) = JAC(c, d) * sq;
        for (int k = 0; k < D; k++) {
            float vk = SF(nt_vec, k);
            if (!(fabsf(vk) > m.eps)) continue;
            float Lkk = SF(nt_H, k * D + k);
            float t = Lkk * Lkk + sign * vk * vk;
            if (t < m.eps) return true;
            float r = sqrtf(t);
            float cc = r / Lkk, cinv = 1.f / cc, ss = vk / Lkk;
            SF(nt_H, k * D + k) = r;
            for (int i = k + 1; i < D; i++) SF(nt_H, i * D + k) = (SF(nt_H, i * D + k) + ss * SF(nt_vec, i) * sign) * cinv;
            for (int i = k + 1; i < D; i++) SF(nt_vec, i) = SF(nt_vec, i) * cc - ss * SF(nt_H, i * D + k);
        }
    }
    return false;
}

// solver.py:2722-2751 + :1893-1910; the conjugate-gradient solver preconditions with the mass matrix instead
// (solver.py:2737-2747, func_solve_mass_batch)
template <class S>
__device__ inline void dev_gradient(const DModel& m, const S& st, int b) {
    const int D = m.D;
    for (int d = 0; d < D; d++) SF(grad, d) = SF(Ma, d) - SF(force, d) - SF(qfrc_constraint, d);
    if (m.solver_type == RS_SOLVER_CG) {
        dev_solve_mass(m, st, b, FV(st, grad, b), FV(st, Mgrad, b));
        return;
    }
    for (int i = 0; i < D; i++) {
        float cur = SF(grad, i);
        for (int j = 0; j < i; j++) cur = cur - SF(nt_H, i * D + j) * SF(Mgrad, j);
        SF(Mgrad, i) = cur / SF(nt_H, i * D + i);
    }
    for (int ii = 0; ii < D; ii++) {
        int i = D - 1 - ii;
        float cur = SF(Mgrad, i);
        for (int j = i + 1; j < D; j++) cur = cur - SF(nt_H, j * D + i) * SF(Mgrad, j);
        SF(Mgrad, i) = cur / SF(nt_H, i * D + i);
    }
}

struct LsP {
    float alpha, cost, grad, hess;
};

// solver.py:2034-2152
template <class S>
__device__ inline LsP dev_ls_init(const DModel& m, const S& st, int b, int& ls_it) {
    const int D = m.D;
    int nc = SF(n_con, 0), ne = SF(n_con_eq, 0), nef = ne + SF(n_con_fl, 0);
    for (int e = 0; e < m.E; e++)
        for (int i = m.entity_dof_start[e]; i < m.entity_dof_end[e]; i++) {
            float mv = 0.f;
            for (int j = m.entity_dof_start[e]; j < m.entity_dof_end[e]; j++) mv = mv + SF(mass_mat, i * D + j) * SF(search, j);
            SF(mv, i) = mv;
        }
    for (int c = 0; c < nc; c++) {
        float jv = 0.f;
        for (int d = 0; d < D; d++) jv = jv + JAC(c, d) * SF(search, d);
        SF(jv, c) = jv;
    }
    float g1 = 0.f, g2 = 0.f;
    for (int d = 0; d < D; d++) {
        float s = SF(search, d);
        g1 = g1 + (s * SF(Ma, d) - s * SF(force, d));
        g2 = g2 + 0.5f * s * SF(mv, d);
    }
    float gauss = SF(gauss, 0);
    SF(quad_gauss, 0) = gauss;
    SF(quad_gauss, 1) = g1;
    SF(quad_gauss, 2) = g2;
    float t0 = gauss, t1 = g1, t2 = g2, e0 = 0.f, e1 = 0.f, e2 = 0.f;
    for (int c = 0; c < nc; c++) {
        float ja = SF(Jaref, c), jv = SF(jv, c), Dc = SF(efc_D, c);
        float f0 = Dc * (0.5f * ja * ja), f1 = Dc * (jv * ja), f2 = Dc * (0.5f * jv * jv);
        if (c < ne) {
            e0 = e0 + f0; e1 = e1 + f1; e2 = e2 + f2;
            t0 = t0 + f0; t1 = t1 + f1; t2 = t2 + f2;
        } else if (c < nef) {
            float f = SF(efc_frictionloss, c), r = SF(diag, c), rf = r * f;
            bool ln = ja <= -rf, lp = ja >= rf;
            if (ln || lp) {
                f0 = (float)ln * f * (-0.5f * rf - ja) + (float)lp * f * (-0.5f * rf + ja);
                f1 = (float)ln * (-f * jv) + (float)lp * (f * jv);
                f2 = 0.f;
            }
            t0 = t0 + f0; t1 = t1 + f1; t2 = t2 + f2;
        } else {
            float a = (float)(ja < 0.f);
            t0 = t0 + f0 * a; t1 = t1 + f1 * a; t2 = t2 + f2 * a;
        }
    }
    SF(eq_sum, 0) = e0;
    SF(eq_sum, 1) = e1;
    SF(eq_sum, 2) = e2;
    float h = 2.f * t2;
    if (h <= 0.f) h = m.eps;
    ls_it = 1;
    return {0.f, t0, t1, h};
}

// solver.py:2155-2355 (n = 1 or 3 candidate steps in one pass over the rows)
template <class S>
__device__ inline void dev_ls_eval(const DModel& m, const S& st, int b, const float* al, int n, LsP* out,
                                   int& ls_it) {
    int nc = SF(n_con, 0), ne = SF(n_con_eq, 0), nef = ne + SF(n_con_fl, 0);
    float b0 = SF(quad_gauss, 0) + SF(eq_sum, 0), b1 = SF(quad_gauss, 1) + SF(eq_sum, 1),
          b2 = SF(quad_gauss, 2) + SF(eq_sum, 2);
    float t[3][3];
    for (int k = 0; k < 3; k++) {
        t[k][0] = b0;
        t[k][1] = b1;
        t[k][2] = b2;
    }
    for (int c = ne; c < nef; c++) {
        float ja = SF(Jaref, c), jv = SF(jv, c), Dc = SF(efc_D, c), f = SF(efc_frictionloss, c), r = SF(diag, c);
        float f0 = Dc * (0.5f * ja * ja), f1 = Dc * (jv * ja), f2 = Dc * (0.5f * jv * jv), rf = r * f;
        for (int k = 0; k < n; k++) {
            float x = ja + al[k] * jv;
            bool ln = x <= -rf, lp = x >= rf;
            float a0 = f0, a1 = f1, a2 = f2;
            if (ln || lp) {
                a0 = (float)ln * f * (-0.5f * rf - ja) + (float)lp * f * (-0.5f * rf + ja);
                a1 = (float)ln * (-f * jv) + (float)lp * (f * jv);
                a2 = 0.f;
            }
            t[k][0] = t[k][0] + a0; t[k][1] = t[k][1] + a1; t[k][2] = t[k][2] + a2;
        }
    }
    for (int c = nef; c < nc; c++) {
        float ja = SF(Jaref, c), jv = SF(jv, c), Dc = SF(efc_D, c);
        float f0 = Dc * (0.5f * ja * ja), f1 = Dc * (jv * ja), f2 = Dc * (0.5f * jv * jv);
        for (int k = 0; k < n; k++) {
            float a = (float)(ja + al[k] * jv < 0.f);
            t[k][0] = t[k][0] + f0 * a; t[k][1] = t[k][1] + f1 * a; t[k][2] = t[k][2] + f2 * a;
        }
    }
    for (int k = 0; k < n; k++) {
        float a = al[k];
        float h = 2.f * t[k][2];
        if (h <= 0.f) h = m.eps;
        out[k] = {a, a * a * t[k][2] + a * t[k][1] + t[k][0], 2.f * a * t[k][2] + t[k][1], h};
    }
    ls_it += n;
}

__device__ inline int bracket(LsP& p, const LsP* c, float& next) {
    int flag = 0;
    for (int i = 0; i < 3; i++) {
        if (p.grad < 0.f && c[i].grad < 0.f && p.grad < c[i].grad) {
            p = c[i];
            flag = 1;
        } else if (p.grad > 0.f && c[i].grad > 0.f && p.grad > c[i].grad) {
            p = c[i];
            flag = 2;
        }
    }
    next = p.alpha;
    if (flag > 0) next = p.alpha - p.grad / p.hess;
    return flag;
}

// solver.py:2427-2597 func_linesearch_batch
template <class S>
__device__ inline float dev_linesearch(const DModel& m, const S& st, int b) {
    const int D = m.D;
    float sn = 0.f;
    for (int d = 0; d < D; d++) sn = sn + SF(search, d) * SF(search, d);
    sn = sqrtf(sn);
    float gtol = m.tolerance * m.ls_tolerance * sn * (m.meaninertia * (float)max(1, D));
    SF(gtol, 0) = gtol;
    int ls_it = 0, res_code = 0;
    float res = 0.f;
    if (sn < m.eps) {
        SF(ls_result, 0) = 1;
        SF(ls_it, 0) = 0;
        return 0.f;
    }
    LsP p0 = dev_ls_init(m, st, b, ls_it);
    LsP p1;
    float a = p0.alpha - p0.grad / p0.hess;
    dev_ls_eval(m, st, b, &a, 1, &p1, ls_it);
    if (p0.cost < p1.cost) p1 = p0;
    bool done = false;
    if (fabsf(p1.grad) < gtol) {
        res_code = fabsf(p1.alpha) < m.eps ? 2 : 0;
        res = p1.alpha;
        done = true;
    }
    LsP p2 = p1;
    int p2update = 0;
    if (!done) {
        float dir = (float)((p1.grad < 0.f) ? 1 : 0) * 2.f - 1.f;
        while (p1.grad * dir <= -gtol && ls_it < m.ls_iterations) {
            p2 = p1;
            p2update = 1;
            float an = p1.alpha - p1.grad / p1.hess;
            dev_ls_eval(m, st, b, &an, 1, &p1, ls_it);
            if (fabsf(p1.grad) < gtol) {
                res = p1.alpha;
                done = true;
                break;
            }
        }
    }
    if (!done && ls_it >= m.ls_iterations) {
        res_code = 3;
        res = p1.alpha;
        done = true;
    }
    if (!done && !p2update) {
        res_code = 6;
        res = p1.alpha;
        done = true;
    }
    if (!done) {
        float al[3] = {p1.alpha - p1.grad / p1.hess, p1.alpha, (p1.alpha + p2.alpha) * 0.5f};
        while (ls_it < m.ls_iterations) {
            LsP c[3];
            dev_ls_eval(m, st, b, al, 3, c, ls_it);
            float n1 = al[0], n2 = al[1];
            bool found = false;
            float ba = 0.f, bc = 0.f;
            for (int i = 0; i < 3; i++)
                if (fabsf(c[i].grad) < gtol && (!found || c[i].cost < bc)) {
                    ba = c[i].alpha;
                    bc = c[i].cost;
                    found = true;
                }
            if (found) {
                res = ba;
                done = true;
                break;
            }
            int b1 = bracket(p1, c, n1), b2 = bracket(p2, c, n2);
            if (b1 == 0 && b2 == 0) {
                res_code = c[2].cost < p0.cost ? 0 : 7;
                res = al[2];
                done = true;
                break;
            }
            al[0] = n1;
            al[1] = n2;
            al[2] = (p1.alpha + p2.alpha) * 0.5f;
        }
        if (!done) {
            if (p1.cost <= p2.cost && p1.cost < p0.cost) {
                res_code = 4;
                res = p1.alpha;
            } else if (p2.cost <= p1.cost && p2.cost < p0.cost) {
                res_code = 4;
                res = p2.alpha;
            } else {
                res_code = 5;
                res = 0.f;
            }
        }
    }
    SF(ls_result, 0) = res_code;
    SF(ls_it, 0) = ls_it;
    return res;
}

// solver.py:2936-3059 (init), :3063-3139 (Newton iteration), :3164-3190 (body), :3241-3261, :3199-3237
template <class S>
__device__ inline void dev_solve(const DModel& m, const S& st, int b) {
    const int D = m.D;
    int nc = SF(n_con, 0);
    bool ws = nc > 0 && SF(is_warmstart, 0);
    for (int d = 0; d < D; d++) SF(qacc, d) = ws ? SF(qacc_ws, d) : SF(acc_smooth, d);
    for (int d1 = 0; d1 < D; d1++) {
        int e = m.dof_entity[d1];
        float ma = 0.f;
        for (int d2 = m.entity_dof_start[e]; d2 < m.entity_dof_end[e]; d2++) ma = ma + SF(mass_mat, d1 * D + d2) * SF(qacc, d2);
        SF(Ma, d1) = ma;
    }
    for (int c = 0; c < nc; c++) {
        float ja = -SF(aref, c);
        for (int d = 0; d < D; d++) ja = ja + JAC(c, d) * SF(qacc, d);
        SF(Jaref, c) = ja;
    }
    float cost = dev_update_constraint(m, st, b, SF(cost, 0));
    SF(cost, 0) = cost;
    int improved = nc > 0;
    const bool cg = m.solver_type == RS_SOLVER_CG;
    if (!cg) dev_hessian_cholesky(m, st, b);
    dev_gradient(m, st, b);
    for (int d = 0; d < D; d++) SF(search, d) = -SF(Mgrad, d);
    int n_iter = 0;
    if (nc > 0) {
        const float tol = (m.meaninertia * (float)max(1, D)) * m.tolerance;
        for (int it = 0; it < m.iterations; it++) {
            n_iter++;
            float alpha = dev_linesearch(m, st, b);
            if (fabsf(alpha) < m.eps) {
                improved = 0;
            } else {
                for (int d = 0; d < D; d++) {
                    SF(qacc, d) = SF(qacc, d) + SF(search, d) * alpha;
                    SF(Ma, d) = SF(Ma, d) + SF(mv, d) * alpha;
                }
                for (int c = 0; c < nc; c++) SF(Jaref, c) = SF(Jaref, c) + SF(jv, c) * alpha;
                if (cg)
                    for (int d = 0; d < D; d++) {
                        SF(cg_prev_grad, d) = SF(grad, d);
                        SF(cg_prev_Mgrad, d) = SF(Mgrad, d);
                    }
                float prev = SF(cost, 0);
                cost = dev_update_constraint(m, st, b, prev);
                SF(cost, 0) = cost;
                if (!cg && dev_cholesky_incremental(m, st, b)) dev_hessian_cholesky(m, st, b);
                dev_gradient(m, st, b);
                float gn = 0.f;
                for (int d = 0; d < D; d++) gn = gn + SF(grad, d) * SF(grad, d);
                gn = sqrtf(gn);
                improved = gn > tol && (prev - cost) > tol;
                if (improved && !cg) {
                    for (int d = 0; d < D; d++) SF(search, d) = -SF(Mgrad, d);
                } else if (improved) {  // solver.py:2865-2884 Polak-Ribiere direction
                    float beta = 0.f, pg_dot_pMg = 0.f;
                    for (int d = 0; d < D; d++) {
                        beta = beta + SF(grad, d) * (SF(Mgrad, d) - SF(cg_prev_Mgrad, d));
                        pg_dot_pMg = pg_dot_pMg + SF(cg_prev_Mgrad, d) * SF(cg_prev_grad, d);
                    }
                    beta = fmaxf(beta / fmaxf(m.eps, pg_dot_pMg), 0.f);
                    SF(cg_pg_dot_pMg, 0) = pg_dot_pMg;
                    SF(cg_beta, 0) = beta;
                    for (int d = 0; d < D; d++) SF(search, d) = -SF(Mgrad, d) + beta * SF(search, d);
                }
            }
            if (!improved) break;
        }
    } else {
        improved = 0;
    }
    SF(improved, 0) = improved;
    SF(n_iter, 0) = n_iter;
    for (int d = 0; d < D; d++) {
        float qa = SF(qacc, d);
        SF(acc, d) = qa;
        SF(qf_constraint, d) = SF(qfrc_constraint, d);
        SF(force, d) = SF(qf_smooth, d) + SF(qfrc_constraint, d);
        SF(qacc_ws, d) = qa;
        if (isnan(qa)) SF(errbits, 0) = SF(errbits, 0) | RS_ERR_INVALID_FORCE_NAN;
    }
    SF(is_warmstart, 0) = 1;
    int cs = SF(n_con_eq, 0) + SF(n_con_fl, 0);
    int ncont = SF(n_contacts, 0);
    for (int i = 0; i < ncont; i++) {
        f3 nrm = LD3(contact_normal, i);
        float fr = SF(contact_friction, i);
        f3 d1, d2, f = zero3();
        orthogonals(nrm, d1, d2);
        for (int k = 0; k < 4; k++) {
            f3 dd = (float)(2 * (k % 2) - 1) * (k < 2 ? d1 : d2);
            f = f + (dd * fr - nrm) * SF(efc_force, i * 4 + k + cs);
        }
        ST3(contact_force, i, f);
    }
}

#undef JAC

}  // namespace rs
