// Smooth dynamics of one environment (one lane per environment, state in [row][env] planes).
// Restates genesis/engine/solvers/rigid/abd/forward_kinematics.py and forward_dynamics.py (cited per function).
#pragma once
#include "rs_device.h"

namespace rs {

// forward_kinematics.py:580-736 func_forward_kinematics_entity
template <class S>
__device__ inline void dev_forward_kinematics(const DModel& m, const S& st, int b) {
    for (int l = 0; l < m.L; l++) {
        f3 pos = mvec3(m.link_pos, l);
        f4q quat = mquat(m.link_quat, l);
        int par = m.link_parent[l];
        if (par != -1) {
            f3 pp = LD3(links_pos, par);
            f4q pq = LD4(links_quat, par);
            pos = pp + qrot(mvec3(m.link_pos, l), pq);
            quat = qtrans(mquat(m.link_quat, l), pq);
        }
        for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
            int jt = m.joint_type[j], qs = m.joint_q_start[j], ds = m.joint_dof_start[j];
            if (jt == JT_FREE) {
                ST3(xanchor, j, mk3(SF(qpos, qs), SF(qpos, qs + 1), SF(qpos, qs + 2)));
                ST3(xaxis, j, mk3(0.f, 0.f, 1.f));
            } else if (jt != JT_FIXED) {
                f3 axis = mk3(0.f, 0.f, 1.f);
                if (jt == JT_REVOLUTE) axis = mvec3(m.dof_motion_ang, ds);
                else if (jt == JT_PRISMATIC) axis = mvec3(m.dof_motion_vel, ds);
                ST3(xanchor, j, qrot(mvec3(m.joint_pos, j), quat) + pos);
                ST3(xaxis, j, qrot(axis, quat));
            }
            if (jt == JT_FREE) {
                pos = mk3(SF(qpos, qs), SF(qpos, qs + 1), SF(qpos, qs + 2));
                f4q q{SF(qpos, qs + 3), SF(qpos, qs + 4), SF(qpos, qs + 5), SF(qpos, qs + 6)};
                float n = sqrtf(nsq4(q));
                quat = {q.w / n, q.x / n, q.y / n, q.z / n};
                f3 xyz = quat_xyz(quat, m.eps);
                for (int k = 0; k < 3; k++) {
                    SF(dof_pos, ds + k) = pos[k];
                    SF(dof_pos, ds + 3 + k) = xyz[k];
                }
            } else if (jt == JT_SPHERICAL) {
                f4q ql{SF(qpos, qs), SF(qpos, qs + 1), SF(qpos, qs + 2), SF(qpos, qs + 3)};
                f3 xyz = quat_xyz(ql, m.eps);
                for (int k = 0; k < 3; k++) SF(dof_pos, ds + k) = xyz[k];
                quat = qtrans(ql, quat);
                pos = LD3(xanchor, j) - qrot(mvec3(m.joint_pos, j), quat);
            } else if (jt == JT_REVOLUTE) {
                f3 axis = mvec3(m.dof_motion_ang, ds);
                float dp = SF(qpos, qs) - m.qpos0[qs];
                SF(dof_pos, ds) = dp;
                f4q ql = rotvec_quat(axis * dp, m.eps);
                quat = qtrans(ql, quat);
                pos = LD3(xanchor, j) - qrot(mvec3(m.joint_pos, j), quat);
            } else if (jt == JT_PRISMATIC) {
                float dp = SF(qpos, qs) - m.qpos0[qs];
                SF(dof_pos, ds) = dp;
                pos = pos + LD3(xaxis, j) * dp;
            }
        }
        if (!(par == -1 && m.link_is_fixed[l])) {
            ST3(links_pos, l, pos);
            ST4(links_quat, l, quat);
        }
    }
}

// forward_kinematics.py:337-576 func_COM_links_entity
template <class S>
__device__ inline void dev_com_links(const DModel& m, const S& st, int b) {
    for (int e = 0; e < m.E; e++) {
        int l0 = m.entity_link_start[e], l1 = m.entity_link_end[e];
        // root accumulators live in root_COM (as the weighted sum) and cinr_mass (as mass_sum) until normalised
        for (int l = l0; l < l1; l++) {
            ST3(root_COM, l, zero3());
            SF(crb_mass, l) = 0.f;
        }
        for (int l = l0; l < l1; l++) {
            float mass = m.link_inertial_mass[l] + 0.f;
            f3 ip;
            f4q iq;
            tr_pose(mvec3(m.link_inertial_pos, l) + zero3(), mquat(m.link_inertial_quat, l), LD3(links_pos, l),
                    LD4(links_quat, l), ip, iq);
            ST3(i_pos, l, ip);
            ST4(i_quat, l, iq);
            int r = m.link_root[l];
            SF(crb_mass, r) = SF(crb_mass, r) + mass;
            ST3(root_COM, r, LD3(root_COM, r) + mass * ip);
        }
        for (int l = l0; l < l1; l++) {
            int r = m.link_root[l];
            if (l == r) {
                float ms = SF(crb_mass, l);
                if (ms > m.eps) ST3(root_COM, l, LD3(root_COM, l) / ms);
                else ST3(root_COM, l, LD3(i_pos, r));
            }
        }
        for (int l = l0; l < l1; l++) ST3(root_COM, l, LD3(root_COM, m.link_root[l]));
        for (int l = l0; l < l1; l++) {
            f3 ipos = LD3(i_pos, l) - LD3(root_COM, l);
            ST3(i_pos, l, ipos);
            float mass = m.link_inertial_mass[l] + 0.f;
            m33 ci;
            f3 cp;
            inertia_offset(mmat3(m.link_inertial_i, l), mass, ipos, LD4(i_quat, l), m.eps, ci, cp);
            ST9(cinr_inertial, l, ci);
            ST3(cinr_pos, l, cp);
            SF(cinr_mass, l) = mass;
        }
        for (int l = l0; l < l1; l++) {
            if (m.link_dof_end[l] == m.link_dof_start[l]) continue;
            for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
                f3 off = LD3(root_COM, l) - LD3(xanchor, j);
                int jt = m.joint_type[j], ds = m.joint_dof_start[j];
                f3 xa = LD3(xaxis, j);
                if (jt == JT_REVOLUTE) {
                    ST3(cdof_ang, ds, xa);
                    ST3(cdof_vel, ds, cross3(xa, off));
                } else if (jt == JT_PRISMATIC) {
                    ST3(cdof_ang, ds, zero3());
                    ST3(cdof_vel, ds, xa);
                } else if (jt == JT_SPHERICAL || jt == JT_FREE) {
                    int base = ds;
                    if (jt == JT_FREE) {
                        for (int k = 0; k < 3; k++) {
                            f3 e3 = zero3();
                            e3[k] = 1.f;
                            ST3(cdof_ang, ds + k, zero3());
                            ST3(cdof_vel, ds + k, e3);
                        }
                        base = ds + 3;
                    }
                    m33 R = quat_R(LD4(links_quat, l), m.eps);
                    for (int k = 0; k < 3; k++) {
                        f3 row = mk3(R.a[0][k], R.a[1][k], R.a[2][k]);
                        ST3(cdof_ang, base + k, row);
                        ST3(cdof_vel, base + k, cross3(row, off));
                    }
                }
            }
        }
    }
}

// forward_kinematics.py:828-864 func_update_geoms_entity
template <class S>
__device__ inline void dev_update_geoms(const DModel& m, const S& st, int b, bool force_fixed) {
    for (int g = 0; g < m.G; g++) {
        if (!force_fixed && m.geom_is_fixed[g]) continue;
        int l = m.geom_link[g];
        f3 p;
        f4q q;
        tr_pose(mvec3(m.geom_pos, g), mquat(m.geom_quat, g), LD3(links_pos, l), LD4(links_quat, l), p, q);
        ST3(geoms_pos, g, p);
        ST4(geoms_quat, g, q);
    }
}

// forward_kinematics.py:993-1116 func_forward_velocity_entity
template <class S>
__device__ inline void dev_forward_velocity(const DModel& m, const S& st, int b) {
    for (int l = 0; l < m.L; l++) {
        f3 cv = zero3(), ca = zero3();
        int p = m.link_parent[l];
        if (p != -1) {
            cv = LD3(cd_vel, p);
            ca = LD3(cd_ang, p);
        }
        for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
            int jt = m.joint_type[j], ds = m.joint_dof_start[j], de = m.joint_dof_end[j];
            if (jt == JT_FREE) {
                for (int k = 0; k < 3; k++) {
                    float v = SF(vel, ds + k);
                    cv = cv + LD3(cdof_vel, ds + k) * v;
                    ca = ca + LD3(cdof_ang, ds + k) * v;
                }
                for (int k = 0; k < 3; k++) {
                    ST3(cdofd_ang, ds + k, zero3());
                    ST3(cdofd_vel, ds + k, zero3());
                    f3 oa, ov;
                    motion_x_motion(ca, cv, LD3(cdof_ang, ds + 3 + k), LD3(cdof_vel, ds + 3 + k), oa, ov);
                    ST3(cdofd_ang, ds + 3 + k, oa);
                    ST3(cdofd_vel, ds + 3 + k, ov);
                }
                for (int k = 0; k < 3; k++) {
                    float v = SF(vel, ds + 3 + k);
                    cv = cv + LD3(cdof_vel, ds + 3 + k) * v;
                    ca = ca + LD3(cdof_ang, ds + 3 + k) * v;
                }
            } else {
                for (int d = ds; d < de; d++) {
                    f3 oa, ov;
                    motion_x_motion(ca, cv, LD3(cdof_ang, d), LD3(cdof_vel, d), oa, ov);
                    ST3(cdofd_ang, d, oa);
                    ST3(cdofd_vel, d, ov);
                }
                for (int d = ds; d < de; d++) {
                    float v = SF(vel, d);
                    cv = cv + LD3(cdof_vel, d) * v;
                    ca = ca + LD3(cdof_ang, d) * v;
                }
            }
        }
        ST3(cd_vel, l, cv);
        ST3(cd_ang, l, ca);
    }
}

template <class S>
__device__ inline void dev_kinematics(const DModel& m, const S& st, int b, bool force_fixed) {
    dev_forward_kinematics(m, st, b);
    dev_com_links(m, st, b);
    dev_update_geoms(m, st, b, force_fixed);
    dev_forward_velocity(m, st, b);
}

// forward_dynamics.py:292-541 func_compute_mass_matrix
template <class S>
__device__ inline void dev_mass_matrix(const DModel& m, const S& st, int b) {
    const int D = m.D;
    for (int l = 0; l < m.L; l++) {
        ST9(crb_inertial, l, LD9(cinr_inertial, l));
        ST3(crb_pos, l, LD3(cinr_pos, l));
        SF(crb_mass, l) = SF(cinr_mass, l);
    }
    for (int e = 0; e < m.E; e++) {
        int l0 = m.entity_link_start[e], l1 = m.entity_link_end[e];
        for (int i = 0; i < l1 - l0; i++) {
            int l = l1 - 1 - i;
            int p = m.link_parent[l];
            if (p != -1) {
                for (int k = 0; k < 9; k++) SF(crb_inertial, 9 * p + k) = SF(crb_inertial, 9 * p + k) + SF(crb_inertial, 9 * l + k);
                SF(crb_mass, p) = SF(crb_mass, p) + SF(crb_mass, l);
                for (int k = 0; k < 3; k++) SF(crb_pos, 3 * p + k) = SF(crb_pos, 3 * p + k) + SF(crb_pos, 3 * l + k);
            }
        }
    }
    for (int l = 0; l < m.L; l++) {
        int d0 = m.link_dof_start[l], d1 = m.link_dof_end[l];
        if (d0 == d1) continue;
        f3 cp = LD3(crb_pos, l);
        m33 ci = LD9(crb_inertial, l);
        float cm = SF(crb_mass, l);
        for (int d = d0; d < d1; d++) {
            f3 fa, fv;
            inertial_mul(cp, ci, cm, LD3(cdof_vel, d), LD3(cdof_ang, d), fa, fv);
            ST3(f_ang, d, fa);
            ST3(f_vel, d, fv);
        }
    }
    for (int e = 0; e < m.E; e++) {
        int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e];
        for (int i = d0; i < d1; i++) {
            f3 fa = LD3(f_ang, i), fv = LD3(f_vel, i);
            for (int j = d0; j <= i; j++)
                SF(mass_mat, i * D + j) =
                    (dot3(fa, LD3(cdof_ang, j)) + dot3(fv, LD3(cdof_vel, j))) * m.mass_parent_mask[i * D + j];
        }
        for (int i = d0; i < d1; i++)
            for (int j = i + 1; j < d1; j++) SF(mass_mat, i * D + j) = SF(mass_mat, j * D + i);
    }
    for (int d = 0; d < D; d++) {
        float v = SF(mass_mat, d * D + d) + m.dof_armature[d];
        if (m.integrator == RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) {
            v = v + m.dof_damping[d] * m.dt;
            if (SF(ctrl_mode, d) <= CTRL_VELOCITY) v = v - m.dof_act_bias[3 * d + 2] * m.dt;
        }
        SF(mass_mat, d * D + d) = v;
    }
}

// forward_dynamics.py:560-602 func_factor_mass (reverse-order L^T D L) of entity e; with implicit_damping (:575-588,
// called from func_implicit_damping) the diagonal gains damping*dt, and under implicitfast -kv*dt for dofs in position
// or velocity control
template <class S>
__device__ inline void dev_factor_mass_entity(const DModel& m, const S& st, int b, int e, bool implicit_damping) {
    const int D = m.D;
    {
        int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e];
        for (int i = d0; i < d1; i++) {
            for (int j = d0; j <= i; j++) SF(mass_L, i * D + j) = SF(mass_mat, i * D + j);
            if (implicit_damping) {
                SF(mass_L, i * D + i) = SF(mass_L, i * D + i) + m.dof_damping[i] * m.dt;
                if (m.integrator == RS_INTEGRATOR_IMPLICITFAST && SF(ctrl_mode, i) <= CTRL_VELOCITY)
                    SF(mass_L, i * D + i) = SF(mass_L, i * D + i) - m.dof_act_bias[3 * i + 2] * m.dt;
            }
        }
        for (int ii = 0; ii < d1 - d0; ii++) {
            int i = d1 - ii - 1;
            float dinv = 1.f / SF(mass_L, i * D + i);
            SF(mass_D_inv, i) = dinv;
            for (int jj = 0; jj < i - d0; jj++) {
                int j = i - jj - 1;
                float a = SF(mass_L, i * D + j) * dinv;
                for (int k = d0; k <= j; k++) SF(mass_L, j * D + k) = SF(mass_L, j * D + k) - a * SF(mass_L, i * D + k);
                SF(mass_L, i * D + j) = a;
            }
            SF(mass_L, i * D + i) = 1.f;
        }
    }
}

template <class S>
__device__ inline void dev_factor_mass(const DModel& m, const S& st, int b) {
    for (int e = 0; e < m.E; e++) dev_factor_mass_entity(m, st, b, e, false);
}

// forward_dynamics.py:1788-1859 func_implicit_damping (integrators Euler and implicitfast, without MuJoCo
// compatibility): every entity with a damped dof, or under implicitfast a position- or velocity-controlled dof with a
// velocity gain, re-factors M + diag(damping*dt [- kv*dt]) and takes acc = that matrix^-1 force; the others keep acc
template <class S>
__device__ inline void dev_implicit_damping(const DModel& m, const S& st, int b) {
    const int D = m.D;
    for (int e = 0; e < m.E; e++) {
        const int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e];
        bool mask = false;
        for (int d = d0; d < d1; d++) {
            if (m.dof_damping[d] > m.eps) mask = true;
            if (m.integrator != RS_INTEGRATOR_EULER && SF(ctrl_mode, d) <= CTRL_VELOCITY &&
                fabsf(m.dof_act_bias[3 * d + 2]) > m.eps)
                mask = true;
        }
        if (!mask) continue;
        dev_factor_mass_entity(m, st, b, e, true);
        for (int ii = 0; ii < d1 - d0; ii++) {
            int i = d1 - ii - 1;
            float cur = SF(force, i);
            for (int j = i + 1; j < d1; j++) cur = cur - SF(mass_L, j * D + i) * SF(acc, j);
            SF(acc, i) = cur;
        }
        for (int i = d0; i < d1; i++) SF(acc, i) = SF(acc, i) * SF(mass_D_inv, i);
        for (int i = d0; i < d1; i++) {
            float cur = SF(acc, i);
            for (int j = d0; j < i; j++) cur = cur - SF(mass_L, i * D + j) * SF(acc, j);
            SF(acc, i) = cur;
        }
    }
}

// forward_dynamics.py:813-894 func_solve_mass_entity: out = M^-1 vec (vec, out: dof planes)
template <class S>
__device__ inline void dev_solve_mass(const DModel& m, const S& st, int b, FView<float> vec, FView<float> out) {
    const int D = m.D;
    for (int e = 0; e < m.E; e++) {
        int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e];
        for (int ii = 0; ii < d1 - d0; ii++) {
            int i = d1 - ii - 1;
            float cur = vec[i];
            for (int j = i + 1; j < d1; j++) cur = cur - SF(mass_L, j * D + i) * out[j];
            out[i] = cur;
        }
        for (int i = d0; i < d1; i++) out[i] = out[i] * SF(mass_D_inv, i);
        for (int i = d0; i < d1; i++) {
            float cur = out[i];
            for (int j = d0; j < i; j++) cur = cur - SF(mass_L, i * D + j) * out[j];
            out[i] = cur;
        }
    }
}

__device__ inline f4q xyz_quat(f3 e) {
    float ai = 0.5f * e[2], aj = -0.5f * e[1], ak = 0.5f * e[0];
    float si = sinf(ai), sj = sinf(aj), sk = sinf(ak);
    float ci = cosf(ai), cj = cosf(aj), ck = cosf(ak);
    float cc = ci * ck, cs = ci * sk, sc = si * ck, ss = si * sk;
    return {cj * cc + sj * ss, cj * cs - sj * sc, -cj * ss - sj * cc, cj * sc - sj * cs};
}

// forward_dynamics.py:956-1182 func_torque_and_passive_force
template <class S>
__device__ inline void dev_torque_passive(const DModel& m, const S& st, int b) {
    for (int l = 0; l < m.L; l++) {
        int ds = m.link_dof_start[l], de = m.link_dof_end[l];
        if (de == ds) continue;
        int jt = m.joint_type[m.link_joint_start[l]];
        for (int d = ds; d < de; d++) {
            float force = 0.f;
            int mode = SF(ctrl_mode, d);
            cfptr bb = &m.dof_act_bias[3 * d];
            if (mode == CTRL_FORCE) {
                force = SF(ctrl_force, d);
            } else if (mode == CTRL_VELOCITY) {
                force = -bb[2] * (SF(ctrl_vel, d) - SF(vel, d));
            } else if (mode == CTRL_POSITION && !(jt == JT_FREE && d >= ds + 3)) {
                float g = m.dof_act_gain[d], p = SF(dof_pos, d);
                force = g * (SF(ctrl_pos, d) - p) + bb[0] + (g + bb[1]) * p + bb[2] * (SF(vel, d) - SF(ctrl_vel, d));
            }
            SF(qf_applied, d) = clampf(force, m.dof_force_range[2 * d], m.dof_force_range[2 * d + 1]);
        }
        if (jt == JT_FREE && (SF(ctrl_mode, ds + 3) == CTRL_POSITION || SF(ctrl_mode, ds + 4) == CTRL_POSITION ||
                              SF(ctrl_mode, ds + 5) == CTRL_POSITION)) {
            f3 xyz = mk3(SF(dof_pos, ds + 3), SF(dof_pos, ds + 4), SF(dof_pos, ds + 5));
            f3 cxyz = mk3(SF(ctrl_pos, ds + 3), SF(ctrl_pos, ds + 4), SF(ctrl_pos, ds + 5));
            f4q qd = qtrans(xyz_quat(cxyz), qinv(xyz_quat(xyz)));
            f3 rv = mk3(qd.x, qd.y, qd.z);
            float s2 = norm3(rv);
            if (s2 > m.eps) {
                float ang = 2.f * atan2f(s2, fabsf(qd.w));
                float is = ang / s2;
                rv = ((qd.w < 0.f ? -1.f : 1.f) * is) * rv;
            }
            for (int k = 0; k < 3; k++) {
                int d = ds + 3 + k;
                cfptr bb = &m.dof_act_bias[3 * d];
                float g = m.dof_act_gain[d];
                float force = g * rv[k] + bb[0] + (g + bb[1]) * SF(dof_pos, d) + bb[2] * (SF(vel, d) - SF(ctrl_vel, d));
                SF(qf_applied, d) = clampf(force, m.dof_force_range[2 * d], m.dof_force_range[2 * d + 1]);
            }
        }
    }
    for (int d = 0; d < m.D; d++) SF(qf_passive, d) = -m.dof_damping[d] * SF(vel, d);
    for (int l = 0; l < m.L; l++) {
        int ds = m.link_dof_start[l], de = m.link_dof_end[l];
        if (de == ds) continue;
        int jt = m.joint_type[m.link_joint_start[l]];
        if (jt != JT_FREE && jt != JT_FIXED)
            for (int d = ds; d < de; d++) SF(qf_passive, d) = SF(qf_passive, d) + (-SF(dof_pos, d)) * m.dof_stiffness[d];
    }
}

// forward_dynamics.py:1186-1285 func_update_acc
template <class S>
__device__ inline void dev_update_acc(const DModel& m, const S& st, int b, bool cacc) {
    f3 g = mk3(m.gx, m.gy, m.gz);
    for (int e = 0; e < m.E; e++) {
        float gc = 1.f - m.entity_gravity_compensation[e];
        for (int l = m.entity_link_start[e]; l < m.entity_link_end[e]; l++) {
            int p = m.link_parent[l];
            f3 dv, da, cl, ca;
            if (p == -1) {
                dv = (-g) * gc;
                da = cl = ca = zero3();
            } else {
                dv = LD3(cdd_vel, p);
                da = LD3(cdd_ang, p);
                if (cacc) {
                    cl = LD3(cacc_lin, p);
                    ca = LD3(cacc_ang, p);
                }
            }
            for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
                float v = SF(vel, d);
                f3 lv = LD3(cdofd_vel, d) * v, la = LD3(cdofd_ang, d) * v;
                dv = dv + lv;
                da = da + la;
                if (cacc) {
                    float a = SF(acc, d);
                    cl = cl + (lv + LD3(cdof_vel, d) * a);
                    ca = ca + (la + LD3(cdof_ang, d) * a);
                }
            }
            ST3(cdd_vel, l, dv);
            ST3(cdd_ang, l, da);
            if (cacc) {
                ST3(cacc_lin, l, cl);
                ST3(cacc_ang, l, ca);
            }
        }
    }
}

// forward_dynamics.py:1289-1400 func_update_force, :1428-1486 func_bias_force, :1507-1563 func_compute_qacc
template <class S>
__device__ inline void dev_rne_and_qacc(const DModel& m, const S& st, int b) {
    for (int l = 0; l < m.L; l++) {
        m33 I = LD9(cinr_inertial, l);
        f3 cp = LD3(cinr_pos, l), cdv = LD3(cd_vel, l), cda = LD3(cd_ang, l);
        float cm = SF(cinr_mass, l);
        f3 f1a, f1v, f2a, f2v, f3a, f3v;
        inertial_mul(cp, I, cm, LD3(cdd_vel, l), LD3(cdd_ang, l), f1a, f1v);
        inertial_mul(cp, I, cm, cdv, cda, f2a, f2v);
        motion_x_force(cda, cdv, f2a, f2v, f3a, f3v);
        ST3(cfrc_vel, l, f1v + f3v + zero3() + zero3());
        ST3(cfrc_ang, l, f1a + f3a + zero3() + zero3());
    }
    for (int e = 0; e < m.E; e++) {
        int l0 = m.entity_link_start[e], l1 = m.entity_link_end[e];
        for (int i = 0; i < l1 - l0; i++) {
            int l = l1 - 1 - i;
            int p = m.link_parent[l];
            if (p != -1) {
                ST3(cfrc_vel, p, LD3(cfrc_vel, p) + LD3(cfrc_vel, l));
                ST3(cfrc_ang, p, LD3(cfrc_ang, p) + LD3(cfrc_ang, l));
            }
        }
    }
    for (int l = 0; l < m.L; l++) {
        f3 fa = LD3(cfrc_ang, l), fv = LD3(cfrc_vel, l);
        for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
            float bias = dot3(LD3(cdof_ang, d), fa) + dot3(LD3(cdof_vel, d), fv);
            SF(qf_bias, d) = bias;
            float f = SF(qf_passive, d) - bias + SF(qf_applied, d);
            SF(force, d) = f;
            SF(qf_smooth, d) = f;
        }
    }
    dev_solve_mass(m, st, b, FV(st, force, b), FV(st, acc_smooth, b));
    for (int d = 0; d < m.D; d++) SF(acc, d) = SF(acc_smooth, d);
}

// forward_dynamics.py:1567-1707 func_integrate + abd/diff.py:26-55 func_copy_next_to_curr
template <class S>
__device__ inline void dev_integrate(const DModel& m, const S& st, int b) {
    const float dt = m.dt;
    for (int d = 0; d < m.D; d++) SF(vel_next, d) = SF(vel, d) + SF(acc, d) * dt;
    for (int l = 0; l < m.L; l++) {
        int ds = m.link_dof_start[l];
        if (m.link_dof_end[l] == ds) continue;
        int qs = m.link_q_start[l], qe = m.link_q_end[l];
        int jt = m.joint_type[m.link_joint_start[l]];
        if (jt == JT_FREE)
            for (int k = 0; k < 3; k++) SF(qpos_next, qs + k) = SF(qpos, qs + k) + SF(vel_next, ds + k) * dt;
        if (jt == JT_SPHERICAL || jt == JT_FREE) {
            int ro = jt == JT_FREE ? 3 : 0;
            f4q r0{SF(qpos, qs + ro), SF(qpos, qs + ro + 1), SF(qpos, qs + ro + 2), SF(qpos, qs + ro + 3)};
            f3 ang = mk3(SF(vel_next, ds + ro), SF(vel_next, ds + ro + 1), SF(vel_next, ds + ro + 2)) * dt;
            f4q r = qtrans(rotvec_quat(ang, m.eps), r0);
            SF(qpos_next, qs + ro) = r.w;
            SF(qpos_next, qs + ro + 1) = r.x;
            SF(qpos_next, qs + ro + 2) = r.y;
            SF(qpos_next, qs + ro + 3) = r.z;
        } else {
            for (int k = 0; k < qe - qs; k++) SF(qpos_next, qs + k) = SF(qpos, qs + k) + SF(vel_next, ds + k) * dt;
        }
    }
    bool valid = true;
    for (int d = 0; d < m.D; d++) valid &= !isnan(SF(vel_next, d));
    for (int q = 0; q < m.Q; q++) valid &= !isnan(SF(qpos_next, q));
    if (valid) {
        for (int d = 0; d < m.D; d++) SF(vel, d) = SF(vel_next, d);
        for (int q = 0; q < m.Q; q++) SF(qpos, q) = SF(qpos_next, q);
    } else {
        SF(errbits, 0) = SF(errbits, 0) | RS_ERR_INVALID_ACC_NAN;
    }
}

// forward_kinematics.py:1297-1320 kernel_update_geom_aabbs
template <class S>
__device__ inline void dev_aabb(const DModel& m, const S& st, int b, int g) {
    f3 p = LD3(geoms_pos, g);
    f4q q = LD4(geoms_quat, g);
    f3 lo = mk3(INFINITY, INFINITY, INFINITY), hi = mk3(-INFINITY, -INFINITY, -INFINITY);
    for (int c = 0; c < 8; c++) {
        cfptr a = &m.geom_init_aabb[24 * g + 3 * c];
        f3 cp = tr_point(mk3(a[0], a[1], a[2]), p, q);
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], cp[k]);
            hi[k] = fmaxf(hi[k], cp[k]);
        }
    }
    ST3(aabb_min, g, lo);
    ST3(aabb_max, g, hi);
}

}  // namespace rs
