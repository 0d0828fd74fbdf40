// ORACLE (test infrastructure only): smooth dynamics of one environment.
// Restates genesis/engine/solvers/rigid/abd/forward_kinematics.py and forward_dynamics.py (cited per function) for a
// single environment, in the reference's loop order.
#pragma once
#include "oracle_model.h"

namespace orc {

// forward_kinematics.py:580-736 func_forward_kinematics_entity (all entities, in link order)
template <class T> void forward_kinematics(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    for (int i_l = 0; i_l < m.L; i_l++) {
        V3<T> pos = m.vec3(m.link_pos, i_l);
        Q4<T> quat = m.quat4(m.link_quat, i_l);
        int parent = m.link_parent[i_l];
        if (parent != -1) {
            V3<T> pp = E::g3(s.links_pos, parent);
            Q4<T> pq = E::g4(s.links_quat, parent);
            pos = pp + transform_by_quat(m.vec3(m.link_pos, i_l), pq);
            quat = transform_quat_by_quat(m.quat4(m.link_quat, i_l), pq);
        }
        for (int i_j = m.link_joint_start[i_l]; i_j < m.link_joint_end[i_l]; i_j++) {
            int jt = m.joint_type[i_j];
            int qs = m.joint_q_start[i_j];
            int ds = m.joint_dof_start[i_j];
            if (jt == JT_FREE) {
                E::s3(s.xanchor, i_j, v3(s.qpos[qs], s.qpos[qs + 1], s.qpos[qs + 2]));
                E::s3(s.xaxis, i_j, v3(T(0), T(0), T(1)));
            } else if (jt != JT_FIXED) {
                V3<T> axis = v3(T(0), T(0), T(1));
                if (jt == JT_REVOLUTE) axis = m.vec3(m.dof_motion_ang, ds);
                else if (jt == JT_PRISMATIC) axis = m.vec3(m.dof_motion_vel, ds);
                E::s3(s.xanchor, i_j, transform_by_quat(m.vec3(m.joint_pos, i_j), quat) + pos);
                E::s3(s.xaxis, i_j, transform_by_quat(axis, quat));
            }
            if (jt == JT_FREE) {
                pos = v3(s.qpos[qs], s.qpos[qs + 1], s.qpos[qs + 2]);
                Q4<T> q{s.qpos[qs + 3], s.qpos[qs + 4], s.qpos[qs + 5], s.qpos[qs + 6]};
                T n = std::sqrt(norm_sqr(q));
                quat = Q4<T>{q.w / n, q.x / n, q.y / n, q.z / n};
                V3<T> xyz = quat_to_xyz(quat, m.EPS);
                for (int j = 0; j < 3; j++) {
                    s.dof_pos[ds + j] = pos[j];
                    s.dof_pos[ds + 3 + j] = xyz[j];
                }
            } else if (jt == JT_SPHERICAL) {
                Q4<T> qloc{s.qpos[qs], s.qpos[qs + 1], s.qpos[qs + 2], s.qpos[qs + 3]};
                V3<T> xyz = quat_to_xyz(qloc, m.EPS);
                for (int j = 0; j < 3; j++) s.dof_pos[ds + j] = xyz[j];
                quat = transform_quat_by_quat(qloc, quat);
                pos = E::g3(s.xanchor, i_j) - transform_by_quat(m.vec3(m.joint_pos, i_j), quat);
            } else if (jt == JT_REVOLUTE) {
                V3<T> axis = m.vec3(m.dof_motion_ang, ds);
                s.dof_pos[ds] = s.qpos[qs] - m.qpos0[qs];
                Q4<T> qloc = rotvec_to_quat(axis * s.dof_pos[ds], m.EPS);
                quat = transform_quat_by_quat(qloc, quat);
                pos = E::g3(s.xanchor, i_j) - transform_by_quat(m.vec3(m.joint_pos, i_j), quat);
            } else if (jt == JT_PRISMATIC) {
                s.dof_pos[ds] = s.qpos[qs] - m.qpos0[qs];
                pos = pos + E::g3(s.xaxis, i_j) * s.dof_pos[ds];
            }
        }
        if (!(parent == -1 && m.link_is_fixed[i_l])) {
            E::s3(s.links_pos, i_l, pos);
            E::s4(s.links_quat, i_l, quat);
        }
    }
}

// forward_kinematics.py:337-576 func_COM_links_entity (per entity)
template <class T> void com_links(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    for (int i_e = 0; i_e < m.E; i_e++) {
        int l0 = m.entity_link_start[i_e], l1 = m.entity_link_end[i_e];
        std::vector<V3<T>> root_bw(m.L);
        std::vector<T> mass_sum(m.L);
        std::vector<V3<T>> i_pos_bw(m.L);
        for (int i_l = l0; i_l < l1; i_l++) {
            root_bw[i_l] = v3(T(0), T(0), T(0));
            mass_sum[i_l] = 0;
        }
        for (int i_l = l0; i_l < l1; i_l++) {
            T mass = m.link_inertial_mass[i_l] + T(0);
            V3<T> ip;
            Q4<T> iq;
            transform_pos_quat_by_trans_quat(m.vec3(m.link_inertial_pos, i_l) + v3(T(0), T(0), T(0)),
                                             m.quat4(m.link_inertial_quat, i_l), E::g3(s.links_pos, i_l),
                                             E::g4(s.links_quat, i_l), ip, iq);
            i_pos_bw[i_l] = ip;
            E::s4(s.i_quat, i_l, iq);
            int r = m.link_root[i_l];
            mass_sum[r] = mass_sum[r] + mass;
            root_bw[r] = root_bw[r] + mass * ip;
        }
        for (int i_l = l0; i_l < l1; i_l++) {
            int r = m.link_root[i_l];
            if (i_l == r) {
                if (mass_sum[i_l] > m.EPS) E::s3(s.root_COM, i_l, root_bw[i_l] / mass_sum[i_l]);
                else E::s3(s.root_COM, i_l, i_pos_bw[r]);
            }
        }
        for (int i_l = l0; i_l < l1; i_l++) E::s3(s.root_COM, i_l, E::g3(s.root_COM, m.link_root[i_l]));
        for (int i_l = l0; i_l < l1; i_l++) {
            V3<T> ipos = i_pos_bw[i_l] - E::g3(s.root_COM, i_l);
            E::s3(s.i_pos, i_l, ipos);
            M3<T> ci;
            V3<T> cp;
            T mass = m.link_inertial_mass[i_l] + T(0);
            transform_inertia_by_trans_quat(m.mat3(m.link_inertial_i, i_l), mass, ipos, E::g4(s.i_quat, i_l), m.EPS,
                                            ci, cp);
            E::s9(s.cinr_inertial, i_l, ci);
            E::s3(s.cinr_pos, i_l, cp);
            s.cinr_mass[i_l] = mass;
        }
        for (int i_l = l0; i_l < l1; i_l++) {
            if (m.link_dof_end[i_l] == m.link_dof_start[i_l]) continue;
            for (int i_j = m.link_joint_start[i_l]; i_j < m.link_joint_end[i_l]; i_j++) {
                V3<T> off = E::g3(s.root_COM, i_l) - E::g3(s.xanchor, i_j);
                int jt = m.joint_type[i_j];
                int ds = m.joint_dof_start[i_j];
                V3<T> xa = E::g3(s.xaxis, i_j);
                if (jt == JT_REVOLUTE) {
                    E::s3(s.cdof_ang, ds, xa);
                    E::s3(s.cdof_vel, ds, cross(xa, off));
                } else if (jt == JT_PRISMATIC) {
                    E::s3(s.cdof_ang, ds, v3(T(0), T(0), T(0)));
                    E::s3(s.cdof_vel, ds, xa);
                } else if (jt == JT_SPHERICAL || jt == JT_FREE) {
                    int base = ds;
                    if (jt == JT_FREE) {
                        for (int i = 0; i < 3; i++) {
                            V3<T> e = v3(T(0), T(0), T(0));
                            e[i] = T(1);
                            E::s3(s.cdof_ang, ds + i, v3(T(0), T(0), T(0)));
                            E::s3(s.cdof_vel, ds + i, e);
                        }
                        base = ds + 3;
                    }
                    M3<T> R = quat_to_R(E::g4(s.links_quat, i_l), m.EPS);
                    for (int i = 0; i < 3; i++) {
                        V3<T> row = v3(R.m[0][i], R.m[1][i], R.m[2][i]);  // row i of R^T
                        E::s3(s.cdof_ang, base + i, row);
                        E::s3(s.cdof_vel, base + i, cross(row, off));
                    }
                }
            }
        }
    }
}

// forward_kinematics.py:828-864 func_update_geoms_entity
template <class T> void update_geoms(const Model<T>& m, Env<T>& s, bool force_fixed) {
    using E = Env<T>;
    for (int i_g = 0; i_g < m.G; i_g++) {
        if (!force_fixed && m.geom_is_fixed[i_g]) continue;
        int l = m.geom_link[i_g];
        V3<T> p;
        Q4<T> q;
        transform_pos_quat_by_trans_quat(m.vec3(m.geom_pos, i_g), m.quat4(m.geom_quat, i_g), E::g3(s.links_pos, l),
                                         E::g4(s.links_quat, l), p, q);
        E::s3(s.geoms_pos, i_g, p);
        E::s4(s.geoms_quat, i_g, q);
    }
}

// forward_kinematics.py:993-1116 func_forward_velocity_entity
template <class T> void forward_velocity(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    for (int i_l = 0; i_l < m.L; i_l++) {
        V3<T> cv = v3(T(0), T(0), T(0)), ca = v3(T(0), T(0), T(0));
        int p = m.link_parent[i_l];
        if (p != -1) {
            cv = E::g3(s.cd_vel, p);
            ca = E::g3(s.cd_ang, p);
        }
        for (int i_j = m.link_joint_start[i_l]; i_j < m.link_joint_end[i_l]; i_j++) {
            int jt = m.joint_type[i_j];
            int ds = m.joint_dof_start[i_j], de = m.joint_dof_end[i_j];
            if (jt == JT_FREE) {
                for (int i = 0; i < 3; i++) {
                    cv = cv + E::g3(s.cdof_vel, ds + i) * s.vel[ds + i];
                    ca = ca + E::g3(s.cdof_ang, ds + i) * s.vel[ds + i];
                }
                for (int i = 0; i < 3; i++) {
                    E::s3(s.cdofd_ang, ds + i, v3(T(0), T(0), T(0)));
                    E::s3(s.cdofd_vel, ds + i, v3(T(0), T(0), T(0)));
                    V3<T> oa, ov;
                    motion_cross_motion(ca, cv, E::g3(s.cdof_ang, ds + 3 + i), E::g3(s.cdof_vel, ds + 3 + i), oa, ov);
                    E::s3(s.cdofd_ang, ds + 3 + i, oa);
                    E::s3(s.cdofd_vel, ds + 3 + i, ov);
                }
                for (int i = 0; i < 3; i++) {
                    cv = cv + E::g3(s.cdof_vel, ds + 3 + i) * s.vel[ds + 3 + i];
                    ca = ca + E::g3(s.cdof_ang, ds + 3 + i) * s.vel[ds + 3 + i];
                }
            } else {
                for (int d = ds; d < de; d++) {
                    V3<T> oa, ov;
                    motion_cross_motion(ca, cv, E::g3(s.cdof_ang, d), E::g3(s.cdof_vel, d), oa, ov);
                    E::s3(s.cdofd_ang, d, oa);
                    E::s3(s.cdofd_vel, d, ov);
                }
                for (int d = ds; d < de; d++) {
                    cv = cv + E::g3(s.cdof_vel, d) * s.vel[d];
                    ca = ca + E::g3(s.cdof_ang, d) * s.vel[d];
                }
            }
        }
        E::s3(s.cd_vel, i_l, cv);
        E::s3(s.cd_ang, i_l, ca);
    }
}

// forward_dynamics.py:292-541 func_compute_mass_matrix (approximate_implicitfast adds damping*dt and kv*dt)
template <class T> void compute_mass_matrix(const Model<T>& m, Env<T>& s, bool implicit_damping) {
    using E = Env<T>;
    const int D = m.D;
    s.crb_inertial = s.cinr_inertial;
    s.crb_pos = s.cinr_pos;
    s.crb_mass = s.cinr_mass;
    for (int i_e = 0; i_e < m.E; i_e++) {
        int l0 = m.entity_link_start[i_e], l1 = m.entity_link_end[i_e];
        for (int i = 0; i < l1 - l0; i++) {
            int i_l = l1 - 1 - i;
            int p = m.link_parent[i_l];
            if (p != -1) {
                for (int k = 0; k < 9; k++) s.crb_inertial[9 * p + k] = s.crb_inertial[9 * p + k] + s.crb_inertial[9 * i_l + k];
                s.crb_mass[p] = s.crb_mass[p] + s.crb_mass[i_l];
                for (int k = 0; k < 3; k++) s.crb_pos[3 * p + k] = s.crb_pos[3 * p + k] + s.crb_pos[3 * i_l + k];
            }
        }
    }
    for (int i_l = 0; i_l < m.L; i_l++) {
        for (int d = m.link_dof_start[i_l]; d < m.link_dof_end[i_l]; d++) {
            V3<T> fa, fv;
            inertial_mul(E::g3(s.crb_pos, i_l), E::g9(s.crb_inertial, i_l), s.crb_mass[i_l], E::g3(s.cdof_vel, d),
                         E::g3(s.cdof_ang, d), fa, fv);
            E::s3(s.f_ang, d, fa);
            E::s3(s.f_vel, d, fv);
        }
    }
    for (int i_e = 0; i_e < m.E; i_e++) {
        int d0 = m.entity_dof_start[i_e], d1 = m.entity_dof_end[i_e];
        for (int i = d0; i < d1; i++)
            for (int j = d0; j < d1; j++)
                s.mass_mat[i * D + j] = (dot(E::g3(s.f_ang, i), E::g3(s.cdof_ang, j)) +
                                         dot(E::g3(s.f_vel, i), E::g3(s.cdof_vel, j))) *
                                        m.mass_parent_mask[i * D + j];
        for (int i = d0; i < d1; i++)
            for (int j = i + 1; j < d1; j++) s.mass_mat[i * D + j] = s.mass_mat[j * D + i];
    }
    for (int d = 0; d < D; d++) s.mass_mat[d * D + d] = s.mass_mat[d * D + d] + m.dof_armature[d];
    if (implicit_damping) {
        T dt = T(m.dt);
        for (int d = 0; d < D; d++) {
            s.mass_mat[d * D + d] = s.mass_mat[d * D + d] + m.dof_damping[d] * dt;
            if (s.ctrl_mode[d] <= CTRL_VELOCITY) s.mass_mat[d * D + d] = s.mass_mat[d * D + d] - m.dof_act_bias[3 * d + 2] * dt;
        }
    }
}

// forward_dynamics.py:560-602 func_factor_mass (serial reverse-order L^T D L) of entity i_e; with implicit_damping
// (:575-588, called from func_implicit_damping) the diagonal gains damping*dt, and under implicitfast -kv*dt for
// dofs in position or velocity control
template <class T> void factor_mass_entity(const Model<T>& m, Env<T>& s, int i_e, bool implicit_damping) {
    const int D = m.D;
    {
        int d0 = m.entity_dof_start[i_e], d1 = m.entity_dof_end[i_e];
        int nd = d1 - d0;
        for (int i = d0; i < d1; i++) {
            for (int j = d0; j <= i; j++) s.mass_L[i * D + j] = s.mass_mat[i * D + j];
            if (implicit_damping) {
                const T dt = T(m.dt);
                s.mass_L[i * D + i] = s.mass_L[i * D + i] + m.dof_damping[i] * dt;
                if (m.h.integrator == RS_INTEGRATOR_IMPLICITFAST && s.ctrl_mode[i] <= CTRL_VELOCITY)
                    s.mass_L[i * D + i] = s.mass_L[i * D + i] - m.dof_act_bias[3 * i + 2] * dt;
            }
        }
        for (int ii = 0; ii < nd; ii++) {
            int i = d1 - ii - 1;
            T D_inv = T(1) / s.mass_L[i * D + i];
            s.mass_D_inv[i] = D_inv;
            for (int jj = 0; jj < i - d0; jj++) {
                int j = i - jj - 1;
                T a = s.mass_L[i * D + j] * D_inv;
                for (int k = d0; k <= j; k++) s.mass_L[j * D + k] -= a * s.mass_L[i * D + k];
                s.mass_L[i * D + j] = a;
            }
            s.mass_L[i * D + i] = T(1);
        }
    }
}

template <class T> void factor_mass(const Model<T>& m, Env<T>& s) {
    for (int i_e = 0; i_e < m.E; i_e++) factor_mass_entity(m, s, i_e, false);
}

// forward_dynamics.py:813-894 func_solve_mass_entity
template <class T> void solve_mass(const Model<T>& m, const Env<T>& s, const std::vector<T>& vec, std::vector<T>& out) {
    const int D = m.D;
    for (int i_e = 0; i_e < m.E; i_e++) {
        int d0 = m.entity_dof_start[i_e], d1 = m.entity_dof_end[i_e];
        for (int ii = 0; ii < d1 - d0; ii++) {
            int i = d1 - ii - 1;
            T cur = vec[i];
            for (int j = i + 1; j < d1; j++) cur = cur - s.mass_L[j * D + i] * out[j];
            out[i] = cur;
        }
        for (int i = d0; i < d1; i++) out[i] = out[i] * s.mass_D_inv[i];
        for (int i = d0; i < d1; i++) {
            T cur = out[i];
            for (int j = d0; j < i; j++) cur = cur - s.mass_L[i * D + j] * out[j];
            out[i] = cur;
        }
    }
}

// forward_dynamics.py:1788-1859 func_implicit_damping (integrators Euler and implicitfast, without MuJoCo
// compatibility): every entity with a damped dof, or under implicitfast a position- or velocity-controlled dof with a
// velocity gain, re-factors M + diag(damping*dt [- kv*dt]) and takes acc = that matrix^-1 force; the others keep acc
template <class T> void implicit_damping(const Model<T>& m, Env<T>& s) {
    const int D = m.D;
    for (int i_e = 0; i_e < m.E; i_e++) {
        const int d0 = m.entity_dof_start[i_e], d1 = m.entity_dof_end[i_e];
        bool mask = false;
        for (int d = d0; d < d1; d++) {
            if (m.dof_damping[d] > m.EPS) mask = true;
            if (m.h.integrator != RS_INTEGRATOR_EULER && s.ctrl_mode[d] <= CTRL_VELOCITY &&
                std::fabs(m.dof_act_bias[3 * d + 2]) > m.EPS)
                mask = true;
        }
        if (!mask) continue;
        factor_mass_entity(m, s, i_e, true);
        for (int ii = 0; ii < d1 - d0; ii++) {
            int i = d1 - ii - 1;
            T cur = s.force[i];
            for (int j = i + 1; j < d1; j++) cur = cur - s.mass_L[j * D + i] * s.acc[j];
            s.acc[i] = cur;
        }
        for (int i = d0; i < d1; i++) s.acc[i] = s.acc[i] * s.mass_D_inv[i];
        for (int i = d0; i < d1; i++) {
            T cur = s.acc[i];
            for (int j = d0; j < i; j++) cur = cur - s.mass_L[i * D + j] * s.acc[j];
            s.acc[i] = cur;
        }
    }
}

// forward_dynamics.py:956-1182 func_torque_and_passive_force
template <class T> void torque_and_passive(const Model<T>& m, Env<T>& s) {
    for (int i_l = 0; i_l < m.L; i_l++) {
        int ds = m.link_dof_start[i_l], de = m.link_dof_end[i_l];
        if (de == ds) continue;
        int jt = m.joint_type[m.link_joint_start[i_l]];
        for (int d = ds; d < de; d++) {
            T force = 0;
            int mode = s.ctrl_mode[d];
            const T* b = &m.dof_act_bias[3 * d];
            if (mode == CTRL_FORCE) {
                force = s.ctrl_force[d];
            } else if (mode == CTRL_VELOCITY) {
                force = -b[2] * (s.ctrl_vel[d] - s.vel[d]);
            } else if (mode == CTRL_POSITION && !(jt == JT_FREE && d >= ds + 3)) {
                T g = m.dof_act_gain[d];
                force = g * (s.ctrl_pos[d] - s.dof_pos[d]) + b[0] + (g + b[1]) * s.dof_pos[d] +
                        b[2] * (s.vel[d] - s.ctrl_vel[d]);
            }
            s.qf_applied[d] = clampv(force, m.dof_force_range[2 * d], m.dof_force_range[2 * d + 1]);
        }
        if (jt == JT_FREE && (s.ctrl_mode[ds + 3] == CTRL_POSITION || s.ctrl_mode[ds + 4] == CTRL_POSITION ||
                              s.ctrl_mode[ds + 5] == CTRL_POSITION)) {
            // Rotational PD of free joints through Euler angles (forward_dynamics.py:1031-1076).
            V3<T> xyz = v3(s.dof_pos[ds + 3], s.dof_pos[ds + 4], s.dof_pos[ds + 5]);
            V3<T> cxyz = v3(s.ctrl_pos[ds + 3], s.ctrl_pos[ds + 4], s.ctrl_pos[ds + 5]);
            auto xyz_to_quat = [](V3<T> e) {
                T ai = T(0.5) * e[2], aj = T(-0.5) * e[1], ak = T(0.5) * e[0];
                T si = std::sin(ai), sj = std::sin(aj), sk = std::sin(ak);
                T ci = std::cos(ai), cj = std::cos(aj), ck = std::cos(ak);
                T cc = ci * ck, cs = ci * sk, sc = si * ck, ss = si * sk;
                return Q4<T>{cj * cc + sj * ss, cj * cs - sj * sc, -cj * ss - sj * cc, cj * sc - sj * cs};
            };
            Q4<T> qd = transform_quat_by_quat(xyz_to_quat(cxyz), inv_quat(xyz_to_quat(xyz)));
            V3<T> rv = v3(qd.x, qd.y, qd.z);
            T s2 = norm(rv);
            if (s2 > m.EPS) {
                T angle = T(2) * std::atan2(s2, std::fabs(qd.w));
                T inv_sinc = angle / s2;
                rv = ((qd.w < T(0) ? T(-1) : T(1)) * inv_sinc) * rv;
            }
            for (int j = 0; j < 3; j++) {
                int d = ds + 3 + j;
                const T* b = &m.dof_act_bias[3 * d];
                T g = m.dof_act_gain[d];
                T force = g * rv[j] + b[0] + (g + b[1]) * s.dof_pos[d] + b[2] * (s.vel[d] - s.ctrl_vel[d]);
                s.qf_applied[d] = clampv(force, m.dof_force_range[2 * d], m.dof_force_range[2 * d + 1]);
            }
        }
    }
    for (int d = 0; d < m.D; d++) s.qf_passive[d] = -m.dof_damping[d] * s.vel[d];
    for (int i_l = 0; i_l < m.L; i_l++) {
        int ds = m.link_dof_start[i_l], de = m.link_dof_end[i_l];
        if (de == ds) continue;
        int jt = m.joint_type[m.link_joint_start[i_l]];
        if (jt != JT_FREE && jt != JT_FIXED)
            for (int d = ds; d < de; d++) s.qf_passive[d] = s.qf_passive[d] + (-s.dof_pos[d]) * m.dof_stiffness[d];
    }
}

// forward_dynamics.py:1186-1285 func_update_acc
template <class T> void update_acc(const Model<T>& m, Env<T>& s, bool update_cacc) {
    using E = Env<T>;
    V3<T> g = v3(T(m.gravity[0]), T(m.gravity[1]), T(m.gravity[2]));
    for (int i_e = 0; i_e < m.E; i_e++) {
        T gc = T(1) - m.entity_gravity_compensation[i_e];
        for (int i_l = m.entity_link_start[i_e]; i_l < m.entity_link_end[i_e]; i_l++) {
            int p = m.link_parent[i_l];
            V3<T> dv, da, cl, ca;
            if (p == -1) {
                dv = (-g) * gc;
                da = v3(T(0), T(0), T(0));
                cl = ca = v3(T(0), T(0), T(0));
            } else {
                dv = E::g3(s.cdd_vel, p);
                da = E::g3(s.cdd_ang, p);
                cl = E::g3(s.cacc_lin, p);
                ca = E::g3(s.cacc_ang, p);
            }
            for (int d = m.link_dof_start[i_l]; d < m.link_dof_end[i_l]; d++) {
                V3<T> lv = E::g3(s.cdofd_vel, d) * s.vel[d];
                V3<T> la = E::g3(s.cdofd_ang, d) * s.vel[d];
                dv = dv + lv;
                da = da + la;
                if (update_cacc) {
                    cl = cl + (lv + E::g3(s.cdof_vel, d) * s.acc[d]);
                    ca = ca + (la + E::g3(s.cdof_ang, d) * s.acc[d]);
                }
            }
            E::s3(s.cdd_vel, i_l, dv);
            E::s3(s.cdd_ang, i_l, da);
            if (update_cacc) {
                E::s3(s.cacc_lin, i_l, cl);
                E::s3(s.cacc_ang, i_l, ca);
            }
        }
    }
}

// forward_dynamics.py:1289-1400 func_update_force
template <class T> void update_force(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    for (int i_l = 0; i_l < m.L; i_l++) {
        V3<T> f1a, f1v, f2a, f2v, f3a, f3v;
        M3<T> I = E::g9(s.cinr_inertial, i_l);
        V3<T> cp = E::g3(s.cinr_pos, i_l);
        T cm = s.cinr_mass[i_l];
        inertial_mul(cp, I, cm, E::g3(s.cdd_vel, i_l), E::g3(s.cdd_ang, i_l), f1a, f1v);
        inertial_mul(cp, I, cm, E::g3(s.cd_vel, i_l), E::g3(s.cd_ang, i_l), f2a, f2v);
        motion_cross_force(E::g3(s.cd_ang, i_l), E::g3(s.cd_vel, i_l), f2a, f2v, f3a, f3v);
        V3<T> z = v3(T(0), T(0), T(0));
        E::s3(s.cfrc_vel, i_l, f1v + f3v + z + z);
        E::s3(s.cfrc_ang, i_l, f1a + f3a + z + z);
    }
    for (int i_e = 0; i_e < m.E; i_e++) {
        int l0 = m.entity_link_start[i_e], l1 = m.entity_link_end[i_e];
        for (int i = 0; i < l1 - l0; i++) {
            int i_l = l1 - 1 - i;
            int p = m.link_parent[i_l];
            if (p != -1) {
                E::s3(s.cfrc_vel, p, E::g3(s.cfrc_vel, p) + E::g3(s.cfrc_vel, i_l));
                E::s3(s.cfrc_ang, p, E::g3(s.cfrc_ang, p) + E::g3(s.cfrc_ang, i_l));
            }
        }
    }
}

// forward_dynamics.py:1428-1486 func_bias_force, :1507-1563 func_compute_qacc
template <class T> void bias_force_and_qacc(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    for (int i_l = 0; i_l < m.L; i_l++)
        for (int d = m.link_dof_start[i_l]; d < m.link_dof_end[i_l]; d++) {
            s.qf_bias[d] = dot(E::g3(s.cdof_ang, d), E::g3(s.cfrc_ang, i_l)) +
                           dot(E::g3(s.cdof_vel, d), E::g3(s.cfrc_vel, i_l));
            s.force[d] = s.qf_passive[d] - s.qf_bias[d] + s.qf_applied[d];
            s.qf_smooth[d] = s.force[d];
        }
    solve_mass(m, s, s.force, s.acc_smooth);
    for (int d = 0; d < m.D; d++) s.acc[d] = s.acc_smooth[d];
}

// forward_kinematics.py:1297-1320 kernel_update_geom_aabbs
template <class T> void update_aabbs(const Model<T>& m, Env<T>& s) {
    using E = Env<T>;
    for (int i_g = 0; i_g < m.G; i_g++) {
        V3<T> p = E::g3(s.geoms_pos, i_g);
        Q4<T> q = E::g4(s.geoms_quat, i_g);
        V3<T> lo = v3(T(INFINITY), T(INFINITY), T(INFINITY)), hi = v3(T(-INFINITY), T(-INFINITY), T(-INFINITY));
        for (int c = 0; c < 8; c++) {
            const T* a = &m.geom_init_aabb[24 * i_g + 3 * c];
            V3<T> cp = transform_by_trans_quat(v3(a[0], a[1], a[2]), p, q);
            for (int k = 0; k < 3; k++) {
                lo[k] = std::fmin(lo[k], cp[k]);
                hi[k] = std::fmax(hi[k], cp[k]);
            }
        }
        E::s3(s.aabb_min, i_g, lo);
        E::s3(s.aabb_max, i_g, hi);
    }
}

// forward_dynamics.py:1567-1707 func_integrate + abd/diff.py:26-55 func_copy_next_to_curr
template <class T> void integrate(const Model<T>& m, Env<T>& s) {
    T dt = T(m.dt);
    std::vector<T> vel_next(m.D), qpos_next(m.Q);
    for (int d = 0; d < m.D; d++) vel_next[d] = s.vel[d] + s.acc[d] * dt;
    for (int i_l = 0; i_l < m.L; i_l++) {
        int ds = m.link_dof_start[i_l];
        if (m.link_dof_end[i_l] == ds) continue;
        int qs = m.link_q_start[i_l], qe = m.link_q_end[i_l];
        int jt = m.joint_type[m.link_joint_start[i_l]];
        if (jt == JT_FREE) {
            for (int j = 0; j < 3; j++) qpos_next[qs + j] = s.qpos[qs + j] + vel_next[ds + j] * dt;
        }
        if (jt == JT_SPHERICAL || jt == JT_FREE) {
            int ro = jt == JT_FREE ? 3 : 0;
            Q4<T> rot0{s.qpos[qs + ro], s.qpos[qs + ro + 1], s.qpos[qs + ro + 2], s.qpos[qs + ro + 3]};
            V3<T> ang = v3(vel_next[ds + ro], vel_next[ds + ro + 1], vel_next[ds + ro + 2]) * dt;
            Q4<T> qrot = rotvec_to_quat(ang, m.EPS);
            Q4<T> rot = transform_quat_by_quat(qrot, rot0);
            for (int j = 0; j < 4; j++) qpos_next[qs + ro + j] = rot[j];
        } else {
            for (int j = 0; j < qe - qs; j++) qpos_next[qs + j] = s.qpos[qs + j] + vel_next[ds + j] * dt;
        }
    }
    bool valid = true;
    for (int d = 0; d < m.D; d++) valid &= !std::isnan(vel_next[d]);
    for (int q = 0; q < m.Q; q++) valid &= !std::isnan(qpos_next[q]);
    if (valid) {
        s.vel = vel_next;
        s.qpos = qpos_next;
    } else {
        s.errno_ |= RS_ERR_INVALID_ACC_NAN;
    }
}

}  // namespace orc
