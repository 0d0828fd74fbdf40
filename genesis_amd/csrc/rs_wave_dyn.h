// Smooth dynamics and kinematics of one environment on one wavefront (fused substep).
// Same arithmetic, in the same order per value, as rs_dyn.h (abd/forward_kinematics.py, abd/forward_dynamics.py):
// per-link, per-dof and per-entry work is split across lanes; accumulations along the kinematic tree (subtree
// inertias, bias forces, spatial velocities and accelerations) are split by vector component, so that each lane walks
// the tree in the reference's order for one component and no lane waits for another.
#pragma once
#include "rs_dyn.h"
#include "rs_wave.h"

namespace rs {

#define LF(field, row) FP(loc, field)[(row)]

// forward_dynamics.py:292-541 func_compute_mass_matrix
__device__ inline void wave_mass_matrix(const DModel& m, const LState& loc, int lane) {
    const int D = m.D, L = m.L;
    for (int t = lane; t < L * 13; t += kWave) {
        const int l = t / 13, k = t - 13 * (t / 13);
        if (k < 9) LF(crb_inertial, 9 * l + k) = LF(cinr_inertial, 9 * l + k);
        else if (k < 12) LF(crb_pos, 3 * l + k - 9) = LF(cinr_pos, 3 * l + k - 9);
        else LF(crb_mass, l) = LF(cinr_mass, l);
    }
    wave_sync();
    // subtree accumulation, leaves to root: lane k owns component k of every link
    if (lane < 13) {
        float* f = lane < 9 ? FP(loc, crb_inertial) : (lane < 12 ? FP(loc, crb_pos) : FP(loc, crb_mass));
        const int w = lane < 9 ? 9 : (lane < 12 ? 3 : 1), c = lane < 9 ? lane : (lane < 12 ? lane - 9 : 0);
        for (int e = 0; e < m.E; e++) {
            const int l0 = m.entity_link_start[e], l1 = m.entity_link_end[e];
            for (int i = 0; i < l1 - l0; i++) {
                const int l = l1 - 1 - i, p = m.link_parent[l];
                if (p != -1) f[w * p + c] = f[w * p + c] + f[w * l + c];
            }
        }
    }
    wave_sync();
    for (int d = lane; d < D; d += kWave) {
        const int l = m.dof_link[d];
        f3 fa, fv;
        inertial_mul(ld3(FP(loc, crb_pos), l, 1, 0), ld9(FP(loc, crb_inertial), l, 1, 0), LF(crb_mass, l),
                     ld3(FP(loc, cdof_vel), d, 1, 0), ld3(FP(loc, cdof_ang), d, 1, 0), fa, fv);
        st3(FP(loc, f_ang), d, 1, 0, fa);
        st3(FP(loc, f_vel), d, 1, 0, fv);
    }
    wave_sync();
    // the entries of each entity's diagonal block (the others stay zero)
    for (int e = 0; e < m.E; e++) {
      const int d0 = m.entity_dof_start[e], n = m.entity_dof_end[e] - d0;
      for (int u = lane; u < n * n; u += kWave) {
        const int i = d0 + u / n, j = d0 + (u - n * (u / n)), t = i * D + j;
        const int a = i >= j ? i : j, c = i >= j ? j : i;  // the lower-triangle entry (a, c) carries the value
        float v = (dot3(ld3(FP(loc, f_ang), a, 1, 0), ld3(FP(loc, cdof_ang), c, 1, 0)) +
                   dot3(ld3(FP(loc, f_vel), a, 1, 0), ld3(FP(loc, cdof_vel), c, 1, 0))) *
                  m.mass_parent_mask[a * D + c];
        if (i == j) {
            v = v + m.dof_armature[i];
            if (m.integrator == RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) {
                v = v + m.dof_damping[i] * m.dt;
                if (LF(ctrl_mode, i) <= CTRL_VELOCITY) v = v - m.dof_act_bias[3 * i + 2] * m.dt;
            }
        }
        LF(mass_mat, t) = v;
      }
    }
    wave_sync();
}

// forward_dynamics.py:560-602 func_factor_mass: for each pivot i (last to first), the rows j < i are updated by
// separate lanes; row i's entries are read before any lane rescales them
__device__ inline void wave_factor_mass(const DModel& m, const LState& loc, int lane) {
    const int D = m.D;
    if (m.E > kWave / 8) {
        // many entities (a pile of free bodies): one entity per lane, each eliminated sequentially; every entry gets
        // the same updates in the same order as below, so the factor is the same to the bit
        for (int e = lane; e < m.E; e += kWave) {
            const int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e];
            for (int i = d0; i < d1; i++)
                for (int j = d0; j <= i; j++) LF(mass_L, i * D + j) = LF(mass_mat, i * D + j);
            for (int ii = 0; ii < d1 - d0; ii++) {
                const int i = d1 - ii - 1;
                const float dinv = 1.f / LF(mass_L, i * D + i);
                LF(mass_D_inv, i) = dinv;
                for (int jj = 0; jj < i - d0; jj++) {
                    const int j = i - jj - 1;
                    const float a = LF(mass_L, i * D + j) * dinv;
                    for (int k = d0; k <= j; k++) LF(mass_L, j * D + k) = LF(mass_L, j * D + k) - a * LF(mass_L, i * D + k);
                    LF(mass_L, i * D + j) = a;
                }
                LF(mass_L, i * D + i) = 1.f;
            }
        }
        wave_sync();
        return;
    }
    for (int e = 0; e < m.E; e++) {
        const int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e], n = d1 - d0;
        for (int t = lane; t < n * n; t += kWave) {
            const int i = d0 + t / n, j = d0 + (t - n * (t / n));
            if (j <= i) LF(mass_L, i * D + j) = LF(mass_mat, i * D + j);
        }
        wave_sync();
        if (n > kWave + 1) {  // more rows than lanes: the sequential elimination, uniformly
            for (int ii = 0; ii < n; ii++) {
                const int i = d1 - ii - 1;
                const float dinv = 1.f / LF(mass_L, i * D + i);
                LF(mass_D_inv, i) = dinv;
                for (int jj = 0; jj < i - d0; jj++) {
                    const int j = i - jj - 1;
                    const float a = LF(mass_L, i * D + j) * dinv;
                    for (int k = d0; k <= j; k++) LF(mass_L, j * D + k) = LF(mass_L, j * D + k) - a * LF(mass_L, i * D + k);
                    LF(mass_L, i * D + j) = a;
                }
                LF(mass_L, i * D + i) = 1.f;
            }
            wave_sync();
            continue;
        }
        for (int ii = 0; ii < n; ii++) {
            const int i = d1 - ii - 1;
            const float dinv = 1.f / LF(mass_L, i * D + i);
            float a = 0.f;
            const int j = d0 + lane;
            if (j < i) {
                a = LF(mass_L, i * D + j) * dinv;
                for (int k = d0; k <= j; k++) LF(mass_L, j * D + k) = LF(mass_L, j * D + k) - a * LF(mass_L, i * D + k);
            }
            wave_sync();
            if (j < i) LF(mass_L, i * D + j) = a;
            LF(mass_D_inv, i) = dinv;
            LF(mass_L, i * D + i) = 1.f;
            wave_sync();
        }
    }
}

// forward_dynamics.py:956-1182 func_torque_and_passive_force, one lane per link (then per dof)
__device__ inline void wave_torque_passive(const DModel& m, const LState& loc, int lane) {
    for (int l = lane; l < m.L; l += kWave) {
        const int ds = m.link_dof_start[l], de = m.link_dof_end[l];
        if (de == ds) continue;
        const int jt = m.joint_type[m.link_joint_start[l]];
        for (int d = ds; d < de; d++) {
            float force = 0.f;
            const int mode = LF(ctrl_mode, d);
            cfptr bb = &m.dof_act_bias[3 * d];
            if (mode == CTRL_FORCE) {
                force = LF(ctrl_force, d);
            } else if (mode == CTRL_VELOCITY) {
                force = -bb[2] * (LF(ctrl_vel, d) - LF(vel, d));
            } else if (mode == CTRL_POSITION && !(jt == JT_FREE && d >= ds + 3)) {
                float g = m.dof_act_gain[d], p = LF(dof_pos, d);
                force = g * (LF(ctrl_pos, d) - p) + bb[0] + (g + bb[1]) * p + bb[2] * (LF(vel, d) - LF(ctrl_vel, d));
            }
            LF(qf_applied, d) = clampf(force, m.dof_force_range[2 * d], m.dof_force_range[2 * d + 1]);
        }
        if (jt == JT_FREE && (LF(ctrl_mode, ds + 3) == CTRL_POSITION || LF(ctrl_mode, ds + 4) == CTRL_POSITION ||
                              LF(ctrl_mode, ds + 5) == CTRL_POSITION)) {
            f3 xyz = mk3(LF(dof_pos, ds + 3), LF(dof_pos, ds + 4), LF(dof_pos, ds + 5));
            f3 cxyz = mk3(LF(ctrl_pos, ds + 3), LF(ctrl_pos, ds + 4), LF(ctrl_pos, ds + 5));
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
                float force = g * rv[k] + bb[0] + (g + bb[1]) * LF(dof_pos, d) + bb[2] * (LF(vel, d) - LF(ctrl_vel, d));
                LF(qf_applied, d) = clampf(force, m.dof_force_range[2 * d], m.dof_force_range[2 * d + 1]);
            }
        }
    }
    for (int d = lane; d < m.D; d += kWave) {
        float p = -m.dof_damping[d] * LF(vel, d);
        const int l = m.dof_link[d];
        const int jt = m.joint_type[m.link_joint_start[l]];
        if (jt != JT_FREE && jt != JT_FIXED) p = p + (-LF(dof_pos, d)) * m.dof_stiffness[d];
        LF(qf_passive, d) = p;
    }
    wave_sync();
}

// forward_dynamics.py:1186-1285 func_update_acc: lane k < 6 owns component k of (cdd_vel, cdd_ang) and, with cacc,
// lane 6 + k of (cacc_lin, cacc_ang), walking the links in order
// dev_solve_mass (M out = vec through the L^T D L factor) with the entities split over the lanes: the mass matrix is
// block-diagonal over entities, and each entity's block is solved by one lane with the same arithmetic as there, so
// the result is the same to the bit.
__device__ inline void wave_solve_mass(const DModel& m, const LState& loc, int lane, const float* vec, float* out) {
    const int D = m.D;
    const float* L = FP(loc, mass_L);
    const float* dinv = FP(loc, mass_D_inv);
    for (int e = lane; e < m.E; e += kWave) {
        const int d0 = m.entity_dof_start[e], d1 = m.entity_dof_end[e];
        for (int ii = 0; ii < d1 - d0; ii++) {
            const int i = d1 - ii - 1;
            float cur = vec[i];
            for (int j = i + 1; j < d1; j++) cur = cur - L[j * D + i] * out[j];
            out[i] = cur;
        }
        for (int i = d0; i < d1; i++) out[i] = out[i] * dinv[i];
        for (int i = d0; i < d1; i++) {
            float cur = out[i];
            for (int j = d0; j < i; j++) cur = cur - L[i * D + j] * out[j];
            out[i] = cur;
        }
    }
    wave_sync();
}

__device__ inline void wave_update_acc(const DModel& m, const LState& loc, int lane, bool cacc) {
    const int nlanes = cacc ? 12 : 6;
    if (lane < nlanes) {
        const int c = lane % 3, grp = lane / 3;  // 0: cdd_vel, 1: cdd_ang, 2: cacc_lin, 3: cacc_ang
        float* out = grp == 0 ? FP(loc, cdd_vel) : grp == 1 ? FP(loc, cdd_ang) : grp == 2 ? FP(loc, cacc_lin) : FP(loc, cacc_ang);
        const float* dd = (grp == 0 || grp == 2) ? FP(loc, cdofd_vel) : FP(loc, cdofd_ang);
        const float* dq = (grp == 0 || grp == 2) ? FP(loc, cdof_vel) : FP(loc, cdof_ang);
        const float g = c == 0 ? m.gx : (c == 1 ? m.gy : m.gz);
        for (int e = 0; e < m.E; e++) {
            const float gc = 1.f - m.entity_gravity_compensation[e];
            for (int l = m.entity_link_start[e]; l < m.entity_link_end[e]; l++) {
                const int p = m.link_parent[l];
                float x = p == -1 ? (grp == 0 ? (-g) * gc : 0.f) : out[3 * p + c];
                for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
                    const float lv = dd[3 * d + c] * LF(vel, d);
                    if (grp < 2) x = x + lv;
                    else x = x + (lv + dq[3 * d + c] * LF(acc, d));
                }
                out[3 * l + c] = x;
            }
        }
    }
    wave_sync();
}

// forward_dynamics.py:1289-1400 func_update_force, :1428-1486 func_bias_force, :1507-1563 func_compute_qacc
__device__ inline void wave_rne_and_qacc(const DModel& m, const LState& loc, int lane) {
    for (int l = lane; l < m.L; l += kWave) {
        m33 I = ld9(FP(loc, cinr_inertial), l, 1, 0);
        f3 cp = ld3(FP(loc, cinr_pos), l, 1, 0), cdv = ld3(FP(loc, cd_vel), l, 1, 0), cda = ld3(FP(loc, cd_ang), l, 1, 0);
        float cm = LF(cinr_mass, l);
        f3 f1a, f1v, f2a, f2v, f3a, f3v;
        inertial_mul(cp, I, cm, ld3(FP(loc, cdd_vel), l, 1, 0), ld3(FP(loc, cdd_ang), l, 1, 0), f1a, f1v);
        inertial_mul(cp, I, cm, cdv, cda, f2a, f2v);
        motion_x_force(cda, cdv, f2a, f2v, f3a, f3v);
        st3(FP(loc, cfrc_vel), l, 1, 0, f1v + f3v + zero3() + zero3());
        st3(FP(loc, cfrc_ang), l, 1, 0, f1a + f3a + zero3() + zero3());
    }
    wave_sync();
    if (lane < 6) {
        float* f = lane < 3 ? FP(loc, cfrc_vel) : FP(loc, cfrc_ang);
        const int c = lane % 3;
        for (int e = 0; e < m.E; e++) {
            const int l0 = m.entity_link_start[e], l1 = m.entity_link_end[e];
            for (int i = 0; i < l1 - l0; i++) {
                const int l = l1 - 1 - i, p = m.link_parent[l];
                if (p != -1) f[3 * p + c] = f[3 * p + c] + f[3 * l + c];
            }
        }
    }
    wave_sync();
    for (int d = lane; d < m.D; d += kWave) {
        const int l = m.dof_link[d];
        f3 fa = ld3(FP(loc, cfrc_ang), l, 1, 0), fv = ld3(FP(loc, cfrc_vel), l, 1, 0);
        float bias = dot3(ld3(FP(loc, cdof_ang), d, 1, 0), fa) + dot3(ld3(FP(loc, cdof_vel), d, 1, 0), fv);
        LF(qf_bias, d) = bias;
        float f = LF(qf_passive, d) - bias + LF(qf_applied, d);
        LF(force, d) = f;
        LF(qf_smooth, d) = f;
    }
    wave_sync();
    wave_solve_mass(m, loc, lane, FP(loc, force), FP(loc, acc_smooth));
    for (int d = lane; d < m.D; d += kWave) LF(acc, d) = LF(acc_smooth, d);
    wave_sync();
}

// forward_kinematics.py:337-576 func_COM_links_entity
__device__ inline void wave_com_links(const DModel& m, const LState& loc, int lane) {
    for (int e = 0; e < m.E; e++) {
        const int l0 = m.entity_link_start[e], l1 = m.entity_link_end[e];
        for (int l = l0 + lane; l < l1; l += kWave) {
            f3 ip;
            f4q iq;
            tr_pose(mvec3(m.link_inertial_pos, l) + zero3(), mquat(m.link_inertial_quat, l), ld3(FP(loc, links_pos), l, 1, 0),
                    ld4(FP(loc, links_quat), l, 1, 0), ip, iq);
            st3(FP(loc, i_pos), l, 1, 0, ip);
            st4(FP(loc, i_quat), l, 1, 0, iq);
            st3(FP(loc, root_COM), l, 1, 0, zero3());
            LF(crb_mass, l) = 0.f;
        }
        wave_sync();
        // root sums in link order: lane 0 the mass, lanes 1..3 the weighted position
        if (lane < 4) {
            for (int l = l0; l < l1; l++) {
                const float mass = m.link_inertial_mass[l] + 0.f;
                const int r = m.link_root[l];
                if (lane == 0) LF(crb_mass, r) = LF(crb_mass, r) + mass;
                else LF(root_COM, 3 * r + lane - 1) = LF(root_COM, 3 * r + lane - 1) + mass * LF(i_pos, 3 * l + lane - 1);
            }
        }
        wave_sync();
        for (int l = l0 + lane; l < l1; l += kWave) {
            const int r = m.link_root[l];
            if (l == r) {
                float ms = LF(crb_mass, l);
                if (ms > m.eps) st3(FP(loc, root_COM), l, 1, 0, ld3(FP(loc, root_COM), l, 1, 0) / ms);
                else st3(FP(loc, root_COM), l, 1, 0, ld3(FP(loc, i_pos), r, 1, 0));
            }
        }
        wave_sync();
        for (int l = l0 + lane; l < l1; l += kWave) {
            const f3 rc = ld3(FP(loc, root_COM), m.link_root[l], 1, 0);
            f3 ipos = ld3(FP(loc, i_pos), l, 1, 0) - rc;
            st3(FP(loc, i_pos), l, 1, 0, ipos);
            float mass = m.link_inertial_mass[l] + 0.f;
            m33 ci;
            f3 cp;
            inertia_offset(mmat3(m.link_inertial_i, l), mass, ipos, ld4(FP(loc, i_quat), l, 1, 0), m.eps, ci, cp);
            st9(FP(loc, cinr_inertial), l, 1, 0, ci);
            st3(FP(loc, cinr_pos), l, 1, 0, cp);
            LF(cinr_mass, l) = mass;
        }
        wave_sync();
        for (int l = l0 + lane; l < l1; l += kWave) {
            if (m.link_root[l] != l) st3(FP(loc, root_COM), l, 1, 0, ld3(FP(loc, root_COM), m.link_root[l], 1, 0));
        }
        for (int j = lane; j < m.J; j += kWave) {
            const int l = m.joint_link[j];
            if (l < l0 || l >= l1 || m.link_dof_end[l] == m.link_dof_start[l]) continue;
            f3 off = ld3(FP(loc, root_COM), m.link_root[l], 1, 0) - ld3(FP(loc, xanchor), j, 1, 0);
            const int jt = m.joint_type[j], ds = m.joint_dof_start[j];
            f3 xa = ld3(FP(loc, xaxis), j, 1, 0);
            if (jt == JT_REVOLUTE) {
                st3(FP(loc, cdof_ang), ds, 1, 0, xa);
                st3(FP(loc, cdof_vel), ds, 1, 0, cross3(xa, off));
            } else if (jt == JT_PRISMATIC) {
                st3(FP(loc, cdof_ang), ds, 1, 0, zero3());
                st3(FP(loc, cdof_vel), ds, 1, 0, xa);
            } else if (jt == JT_SPHERICAL || jt == JT_FREE) {
                int base = ds;
                if (jt == JT_FREE) {
                    for (int k = 0; k < 3; k++) {
                        f3 e3 = zero3();
                        e3[k] = 1.f;
                        st3(FP(loc, cdof_ang), ds + k, 1, 0, zero3());
                        st3(FP(loc, cdof_vel), ds + k, 1, 0, e3);
                    }
                    base = ds + 3;
                }
                m33 R = quat_R(ld4(FP(loc, links_quat), l, 1, 0), m.eps);
                for (int k = 0; k < 3; k++) {
                    f3 row = mk3(R.a[0][k], R.a[1][k], R.a[2][k]);
                    st3(FP(loc, cdof_ang), base + k, 1, 0, row);
                    st3(FP(loc, cdof_vel), base + k, 1, 0, cross3(row, off));
                }
            }
        }
        wave_sync();
    }
}

// forward_kinematics.py:993-1116 func_forward_velocity_entity. Lane k < 6 walks the links for component k of
// (cd_vel, cd_ang) and records, for every dof, the (cd_ang, cd_vel) that its cdofd is taken against (the
// velocity accumulated before the dof's joint - after the translational dofs for a free joint's rotations); then one
// lane per dof forms cdofd.
__device__ inline void wave_forward_velocity(const DModel& m, const LState& loc, int lane) {
    float* pre = FP(loc, wave_scratch);  // [d * 6 + (0..2: ang, 3..5: vel)]
    if (lane < 6) {
        const int c = lane % 3;
        const bool ang = lane >= 3;
        float* out = ang ? FP(loc, cd_ang) : FP(loc, cd_vel);
        const float* cd = ang ? FP(loc, cdof_ang) : FP(loc, cdof_vel);
        const int slot = ang ? c : 3 + c;
        for (int l = 0; l < m.L; l++) {
            const int p = m.link_parent[l];
            float x = p != -1 ? out[3 * p + c] : 0.f;
            for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
                const int jt = m.joint_type[j], ds = m.joint_dof_start[j], de = m.joint_dof_end[j];
                if (jt == JT_FREE) {
                    for (int k = 0; k < 3; k++) x = x + cd[3 * (ds + k) + c] * LF(vel, ds + k);
                    for (int k = 3; k < 6; k++) pre[6 * (ds + k) + slot] = x;
                    for (int k = 3; k < 6; k++) x = x + cd[3 * (ds + k) + c] * LF(vel, ds + k);
                } else {
                    for (int d = ds; d < de; d++) pre[6 * d + slot] = x;
                    for (int d = ds; d < de; d++) x = x + cd[3 * d + c] * LF(vel, d);
                }
            }
            out[3 * l + c] = x;
        }
    }
    wave_sync();
    for (int d = lane; d < m.D; d += kWave) {
        const int l = m.dof_link[d];
        int j = m.link_joint_start[l];
        while (j + 1 < m.link_joint_end[l] && m.joint_dof_end[j] <= d) j++;
        f3 oa = zero3(), ov = zero3();
        if (!(m.joint_type[j] == JT_FREE && d < m.joint_dof_start[j] + 3)) {
            f3 ca = mk3(pre[6 * d], pre[6 * d + 1], pre[6 * d + 2]), cv = mk3(pre[6 * d + 3], pre[6 * d + 4], pre[6 * d + 5]);
            motion_x_motion(ca, cv, ld3(FP(loc, cdof_ang), d, 1, 0), ld3(FP(loc, cdof_vel), d, 1, 0), oa, ov);
        }
        st3(FP(loc, cdofd_ang), d, 1, 0, oa);
        st3(FP(loc, cdofd_vel), d, 1, 0, ov);
    }
    wave_sync();
}

// forward_kinematics.py:828-864 func_update_geoms_entity, one lane per geom
__device__ inline void wave_update_geoms(const DModel& m, const LState& loc, int lane) {
    for (int g = lane; g < m.G; g += kWave) {
        if (m.geom_is_fixed[g]) continue;
        const int l = m.geom_link[g];
        f3 p;
        f4q q;
        tr_pose(mvec3(m.geom_pos, g), mquat(m.geom_quat, g), ld3(FP(loc, links_pos), l, 1, 0), ld4(FP(loc, links_quat), l, 1, 0), p,
                q);
        st3(FP(loc, geoms_pos), g, 1, 0, p);
        st4(FP(loc, geoms_quat), g, 1, 0, q);
    }
}

__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ f3 readlane3(f3 v, int l) { return mk3(readlane_f(v.x, l), readlane_f(v.y, l), readlane_f(v.z, l)); }

// forward_kinematics.py:580-736 with the work that does not depend on the chain done first, one lane per link and
// per joint: the links' and joints' constants (offsets, parents, joint ranges, anchors, axes), the joint's local
// rotation (revolute: from its angle; spherical and free: from its coordinates) and its dof positions. The chain walk
// then runs on every lane with the same arithmetic as dev_forward_kinematics, taking a link's or joint's values from
// its lane (the model reads would otherwise be a chain of dependent loads per link) and keeping the anchor, axis and
// the previous link's pose in registers rather than reading back what it has just stored. Models with more than 64
// links or joints take dev_forward_kinematics.
__device__ inline void wave_forward_kinematics(const DModel& m, const LState& loc, int lane) {
    if (m.J > kWave || m.L > kWave) {
        dev_forward_kinematics(m, loc, 0);
        return;
    }
    const LState& st = loc;
    const int b = 0;
    f4q jq{1.f, 0.f, 0.f, 0.f};  // lane j: local rotation of joint j (free: the normalised orientation)
    float jd = 0.f;               // lane j: revolute / prismatic displacement
    int jtype = JT_FIXED, jqs = 0;
    f3 jpos = zero3(), jaxis = mk3(0.f, 0.f, 1.f);
    if (lane < m.J) {
        const int j = lane;
        const int jt = m.joint_type[j], qs = m.joint_q_start[j], ds = m.joint_dof_start[j];
        jtype = jt;
        jqs = qs;
        jpos = mvec3(m.joint_pos, j);
        if (jt == JT_REVOLUTE) jaxis = mvec3(m.dof_motion_ang, ds);
        else if (jt == JT_PRISMATIC) jaxis = mvec3(m.dof_motion_vel, ds);
        if (jt == JT_FREE) {
            f4q q{SF(qpos, qs + 3), SF(qpos, qs + 4), SF(qpos, qs + 5), SF(qpos, qs + 6)};
            float n = sqrtf(nsq4(q));
            jq = {q.w / n, q.x / n, q.y / n, q.z / n};
            f3 xyz = quat_xyz(jq, m.eps);
            for (int k = 0; k < 3; k++) {
                SF(dof_pos, ds + k) = SF(qpos, qs + k);
                SF(dof_pos, ds + 3 + k) = xyz[k];
            }
        } else if (jt == JT_SPHERICAL) {
            jq = {SF(qpos, qs), SF(qpos, qs + 1), SF(qpos, qs + 2), SF(qpos, qs + 3)};
            f3 xyz = quat_xyz(jq, m.eps);
            for (int k = 0; k < 3; k++) SF(dof_pos, ds + k) = xyz[k];
        } else if (jt == JT_REVOLUTE) {
            jd = SF(qpos, qs) - m.qpos0[qs];
            SF(dof_pos, ds) = jd;
            jq = rotvec_quat(mvec3(m.dof_motion_ang, ds) * jd, m.eps);
        } else if (jt == JT_PRISMATIC) {
            jd = SF(qpos, qs) - m.qpos0[qs];
            SF(dof_pos, ds) = jd;
        }
    }
    f3 lpos = zero3();
    f4q lquat{1.f, 0.f, 0.f, 0.f};
    int lpar = -1, ljs = 0, lje = 0, lfix = 0;
    if (lane < m.L) {
        lpos = mvec3(m.link_pos, lane);
        lquat = mquat(m.link_quat, lane);
        lpar = m.link_parent[lane];
        ljs = m.link_joint_start[lane];
        lje = m.link_joint_end[lane];
        lfix = m.link_is_fixed[lane];
    }
    f3 prev_pos = zero3();
    f4q prev_quat{1.f, 0.f, 0.f, 0.f};
    int prev = -1;  // link whose pose prev_pos / prev_quat hold (stored in this walk)
    for (int l = 0; l < m.L; l++) {
        const f3 lp = readlane3(lpos, l);
        const f4q lq{readlane_f(lquat.w, l), readlane_f(lquat.x, l), readlane_f(lquat.y, l), readlane_f(lquat.z, l)};
        const int par = readlane_i(lpar, l), j0 = readlane_i(ljs, l), j1 = readlane_i(lje, l);
        f3 pos = lp;
        f4q quat = lq;
        if (par != -1) {
            const f3 pp = par == prev ? prev_pos : LD3(links_pos, par);
            const f4q pq = par == prev ? prev_quat : LD4(links_quat, par);
            pos = pp + qrot(lp, pq);
            quat = qtrans(lq, pq);
        }
        for (int j = j0; j < j1; j++) {
            const int jt = readlane_i(jtype, j), qs = readlane_i(jqs, j);
            const f4q ql{readlane_f(jq.w, j), readlane_f(jq.x, j), readlane_f(jq.y, j), readlane_f(jq.z, j)};
            const float dp = readlane_f(jd, j);
            const f3 jp = readlane3(jpos, j);
            f3 xanchor = zero3(), xaxis = zero3();
            if (jt == JT_FREE) {
                xanchor = mk3(SF(qpos, qs), SF(qpos, qs + 1), SF(qpos, qs + 2));
                xaxis = mk3(0.f, 0.f, 1.f);
            } else if (jt != JT_FIXED) {
                xanchor = qrot(jp, quat) + pos;
                xaxis = qrot(readlane3(jaxis, j), quat);
            }
            if (jt != JT_FIXED) {
                ST3(xanchor, j, xanchor);
                ST3(xaxis, j, xaxis);
            }
            if (jt == JT_FREE) {
                pos = xanchor;
                quat = ql;
            } else if (jt == JT_SPHERICAL) {
                quat = qtrans(ql, quat);
                pos = xanchor - qrot(jp, quat);
            } else if (jt == JT_REVOLUTE) {
                quat = qtrans(ql, quat);
                pos = xanchor - qrot(jp, quat);
            } else if (jt == JT_PRISMATIC) {
                pos = pos + xaxis * dp;
            }
        }
        if (!(par == -1 && readlane_i(lfix, l))) {
            ST3(links_pos, l, pos);
            ST4(links_quat, l, quat);
            prev = l;
            prev_pos = pos;
            prev_quat = quat;
        }
    }
    wave_sync();
}

// forward kinematics, then COM, geoms and velocities split across lanes
__device__ inline void wave_kinematics(const DModel& m, const LState& loc, int lane) {
    wave_forward_kinematics(m, loc, lane);
    wave_sync();
    wave_com_links(m, loc, lane);
    wave_update_geoms(m, loc, lane);
    wave_forward_velocity(m, loc, lane);
}

#undef LF

}  // namespace rs
