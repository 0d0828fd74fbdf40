// ORACLE (test infrastructure only): model and per-environment state containers of the CPU restatement.
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/rigidstep.h"
#include "../include/rigidstep_arrays.h"
#include "oracle_math.h"

namespace orc {

constexpr int CTRL_POSITION = 0, CTRL_VELOCITY = 1, CTRL_FORCE = 2;
constexpr int JT_FIXED = 0, JT_REVOLUTE = 1, JT_PRISMATIC = 2, JT_SPHERICAL = 3, JT_FREE = 4;
constexpr int GT_PLANE = 0, GT_SPHERE = 1, GT_ELLIPSOID = 2, GT_CYLINDER = 3, GT_CAPSULE = 4, GT_BOX = 5, GT_MESH = 6,
              GT_TERRAIN = 7;
constexpr int EQ_CONNECT = 0, EQ_WELD = 1, EQ_JOINT = 2;
constexpr int SOLVER_CG = 0, SOLVER_NEWTON = 1;


template <class T> struct Model {
    rs_model_header h;
    T EPS;
    // real-valued options in double (the header carries them as float32 for the device; the FP64 restatement takes
    // the exact values through orc_set_options)
    double dt, gravity[3], tolerance, ls_tolerance, meaninertia, mc_perturbation, mc_tolerance;
    int L, J, D, Q, G, E, NE, P, NSUP;
    int KMAX, CMAX, BROAD;
#define ORC_DECL(name, is_int, kind, width) std::vector<typename std::conditional<is_int, int32_t, T>::type> name;
    RS_MODEL_ARRAYS(ORC_DECL)
#undef ORC_DECL

    V3<T> vec3(const std::vector<T>& a, int i) const { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
    Q4<T> quat4(const std::vector<T>& a, int i) const { return {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}; }
    M3<T> mat3(const std::vector<T>& a, int i) const {
        M3<T> m;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) m.m[r][c] = a[9 * i + 3 * r + c];
        return m;
    }
};

inline int64_t count_of(const rs_model_header& h, int kind) {
    switch (kind) {
        case RS_N_ONE: return 1;
        case RS_N_LINKS: return h.n_links;
        case RS_N_JOINTS: return h.n_joints;
        case RS_N_DOFS: return h.n_dofs;
        case RS_N_QS: return h.n_qs;
        case RS_N_GEOMS: return h.n_geoms;
        case RS_N_ENTITIES: return h.n_entities;
        case RS_N_EQUALITIES: return h.n_equalities;
        case RS_N_DOFS2: return (int64_t)h.n_dofs * h.n_dofs;
        case RS_N_GEOMS2: return (int64_t)h.n_geoms * h.n_geoms;
        case RS_N_PAIRS: return h.n_pairs;
        case RS_N_SUPPORT: return h.n_support;
        case RS_N_VERTS: return h.n_verts;
        default: return 0;
    }
}

// Load from a blob whose floats are float64 (header.float_bytes == 8).
template <class T> bool load_model(Model<T>& m, const rs_model_header* h, const uint8_t* blob) {
    m.h = *h;
    if (h->float_bytes != 8) return false;
    m.EPS = sizeof(T) == 4 ? T(1.1920928955078125e-07) : T(1e-15);
    m.L = h->n_links; m.J = h->n_joints; m.D = h->n_dofs; m.Q = h->n_qs; m.G = h->n_geoms;
    m.E = h->n_entities; m.NE = h->n_equalities; m.P = h->n_pairs; m.NSUP = h->n_support;
    m.KMAX = h->max_contact_pairs; m.CMAX = h->len_constraints; m.BROAD = h->max_collision_pairs_broad;
    m.dt = h->dt;
    for (int k = 0; k < 3; k++) m.gravity[k] = h->gravity[k];
    m.tolerance = h->tolerance;
    m.ls_tolerance = h->ls_tolerance;
    m.meaninertia = h->meaninertia;
    m.mc_perturbation = h->mc_perturbation;
    m.mc_tolerance = h->mc_tolerance;
    int idx = 0;
#define ORC_LOAD(name, is_int, kind, width)                                                   \
    {                                                                                         \
        int64_t n = count_of(*h, kind) * (width);                                             \
        m.name.resize(n);                                                                     \
        const uint8_t* src = blob + h->array_offset[idx];                                     \
        for (int64_t i = 0; i < n; i++) {                                                     \
            if (is_int) {                                                                     \
                int32_t v; std::memcpy(&v, src + 4 * i, 4); m.name[i] = (decltype(m.name[0] + 0))v; \
            } else {                                                                          \
                double v; std::memcpy(&v, src + 8 * i, 8); m.name[i] = (decltype(m.name[0] + 0))v; \
            }                                                                                 \
        }                                                                                     \
        idx++;                                                                                \
    }
    RS_MODEL_ARRAYS(ORC_LOAD)
#undef ORC_LOAD
    return true;
}

// Per-environment state. Vector quantities are stored element-major: [elem * width + comp].
template <class T> struct Env {
    // persistent
    std::vector<T> qpos, vel, acc, ctrl_pos, ctrl_vel, ctrl_force, qacc_ws;
    std::vector<int32_t> ctrl_mode;
    int32_t is_warmstart = 0, errno_ = 0, first_time = 1;
    std::vector<T> sort_value, cache_normal;
    std::vector<int32_t> sort_i_g, sort_is_max;
    // kinematics
    std::vector<T> links_pos, links_quat, root_COM, i_pos, i_quat, cinr_inertial, cinr_pos, cinr_mass, cd_ang, cd_vel;
    std::vector<T> xanchor, xaxis, dof_pos, cdof_ang, cdof_vel, cdofd_ang, cdofd_vel;
    std::vector<T> geoms_pos, geoms_quat, aabb_min, aabb_max;
    // dynamics
    std::vector<T> crb_inertial, crb_pos, crb_mass, f_ang, f_vel, mass_mat, mass_L, mass_D_inv;
    std::vector<T> qf_applied, qf_passive, qf_bias, force, qf_smooth, acc_smooth, qf_constraint;
    std::vector<T> cdd_ang, cdd_vel, cfrc_ang, cfrc_vel, cacc_ang, cacc_lin;
    // collision
    int32_t n_broad = 0, n_contacts = 0;
    // MPR detections this step that the reference would hand to GJK/EPA (narrowphase.py:729-734; not built: the MPR
    // result is kept), first detections and perturbed ones
    int32_t n_gjk_regime = 0, n_gjk_regime_perturbed = 0;
    std::vector<int32_t> broad_pairs, active_buffer;
    std::vector<int32_t> contact_geom_a, contact_geom_b, contact_link_a, contact_link_b, contact_pair;
    std::vector<T> contact_pos, contact_normal, contact_penetration, contact_friction, contact_sol_params,
        contact_force;
    // constraints / solver
    int32_t n_con = 0, n_con_eq = 0, n_con_fl = 0, improved = 0, ls_it = 0, ls_result = 0, n_iter = 0;
    std::vector<T> jac, diag, aref, efc_D, efc_frictionloss, Jaref, jv, efc_force;
    std::vector<int32_t> active, prev_active;
    std::vector<T> qacc, Ma, grad, Mgrad, search, mv, qfrc_constraint, nt_H, nt_vec, cg_prev_grad, cg_prev_Mgrad;
    T cost = 0, prev_cost = 0, gauss = 0, gtol = 0, cg_beta = 0, cg_pg_dot_pMg = 0;
    T quad_gauss[3] = {0, 0, 0}, eq_sum[3] = {0, 0, 0};

    void allocate(const Model<T>& m) {
        int L = m.L, J = m.J, D = m.D, Q = m.Q, G = m.G, P = m.P, K = m.KMAX, C = m.CMAX;
        qpos.assign(Q, 0); vel.assign(D, 0); acc.assign(D, 0); ctrl_pos.assign(D, 0); ctrl_vel.assign(D, 0);
        ctrl_force.assign(D, 0); qacc_ws.assign(D, 0); ctrl_mode.assign(D, CTRL_FORCE);
        sort_value.assign(2 * G, 0); sort_i_g.assign(2 * G, 0); sort_is_max.assign(2 * G, 0);
        cache_normal.assign(3 * P, 0);
        links_pos.assign(3 * L, 0); links_quat.assign(4 * L, 0); root_COM.assign(3 * L, 0); i_pos.assign(3 * L, 0);
        i_quat.assign(4 * L, 0); cinr_inertial.assign(9 * L, 0); cinr_pos.assign(3 * L, 0); cinr_mass.assign(L, 0);
        cd_ang.assign(3 * L, 0); cd_vel.assign(3 * L, 0); xanchor.assign(3 * J, 0); xaxis.assign(3 * J, 0);
        dof_pos.assign(D, 0); cdof_ang.assign(3 * D, 0); cdof_vel.assign(3 * D, 0); cdofd_ang.assign(3 * D, 0);
        cdofd_vel.assign(3 * D, 0); geoms_pos.assign(3 * G, 0); geoms_quat.assign(4 * G, 0);
        aabb_min.assign(3 * G, 0); aabb_max.assign(3 * G, 0);
        crb_inertial.assign(9 * L, 0); crb_pos.assign(3 * L, 0); crb_mass.assign(L, 0); f_ang.assign(3 * D, 0);
        f_vel.assign(3 * D, 0); mass_mat.assign(D * D, 0); mass_L.assign(D * D, 0); mass_D_inv.assign(D, 0);
        qf_applied.assign(D, 0); qf_passive.assign(D, 0); qf_bias.assign(D, 0); force.assign(D, 0);
        qf_smooth.assign(D, 0); acc_smooth.assign(D, 0); qf_constraint.assign(D, 0);
        cdd_ang.assign(3 * L, 0); cdd_vel.assign(3 * L, 0); cfrc_ang.assign(3 * L, 0); cfrc_vel.assign(3 * L, 0);
        cacc_ang.assign(3 * L, 0); cacc_lin.assign(3 * L, 0);
        broad_pairs.assign(2 * std::max(m.BROAD, 1), 0); active_buffer.assign(std::max(G, 1), 0);
        contact_geom_a.assign(K, -1); contact_geom_b.assign(K, -1); contact_link_a.assign(K, -1);
        contact_link_b.assign(K, -1); contact_pair.assign(K, -1); contact_pos.assign(3 * K, 0);
        contact_normal.assign(3 * K, 0); contact_penetration.assign(K, 0); contact_friction.assign(K, 0);
        contact_sol_params.assign(7 * K, 0); contact_force.assign(3 * K, 0);
        jac.assign((size_t)C * D, 0); diag.assign(C, 0); aref.assign(C, 0); efc_D.assign(C, 0);
        efc_frictionloss.assign(C, 0); Jaref.assign(C, 0); jv.assign(C, 0); efc_force.assign(C, 0);
        active.assign(C, 0); prev_active.assign(C, 0);
        qacc.assign(D, 0); Ma.assign(D, 0); grad.assign(D, 0); Mgrad.assign(D, 0); search.assign(D, 0);
        cg_prev_grad.assign(D, 0); cg_prev_Mgrad.assign(D, 0);
        mv.assign(D, 0); qfrc_constraint.assign(D, 0); nt_H.assign(D * D, 0); nt_vec.assign(D, 0);
    }

    static V3<T> g3(const std::vector<T>& a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
    static void s3(std::vector<T>& a, int i, V3<T> v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }
    static Q4<T> g4(const std::vector<T>& a, int i) { return {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}; }
    static void s4(std::vector<T>& a, int i, Q4<T> q) {
        a[4 * i] = q.w; a[4 * i + 1] = q.x; a[4 * i + 2] = q.y; a[4 * i + 3] = q.z;
    }
    static M3<T> g9(const std::vector<T>& a, int i) {
        M3<T> m;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) m.m[r][c] = a[9 * i + 3 * r + c];
        return m;
    }
    static void s9(std::vector<T>& a, int i, const M3<T>& m) {
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) a[9 * i + 3 * r + c] = m.m[r][c];
    }
};

}  // namespace orc
