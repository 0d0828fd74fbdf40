// ORACLE (test infrastructure only). C entry points of the CPU restatement, loaded by tests/ and by bench.py's
// cpu_baseline leg through ctypes. Never linked into the product library.
#include <chrono>
#include <functional>
#include <map>
#include <string>
#include <thread>

#include "oracle_flops.h"  // before the restatement: its std:: overloads must be visible to the templates

#include "oracle_collision.h"
#include "oracle_constraint.h"
#include "oracle_dynamics.h"

namespace orc {

struct FieldRef {
    std::function<double*(void)> unused;
};

template <class T> struct Context {
    Model<T> m;
    std::vector<Env<T>> envs;
    struct Field {
        int is_int;
        std::function<size_t(Env<T>&)> rows;
        std::function<void(Env<T>&, const double*)> set;
        std::function<void(Env<T>&, double*)> get;
    };
    std::map<std::string, Field> fields;

    template <class V> void reg(const char* name, V Env<T>::*mem) {
        Field f;
        f.is_int = std::is_same<typename V::value_type, int32_t>::value;
        f.rows = [mem](Env<T>& e) { return (e.*mem).size(); };
        f.set = [mem](Env<T>& e, const double* in) {
            for (size_t i = 0; i < (e.*mem).size(); i++) (e.*mem)[i] = (typename V::value_type)in[i];
        };
        f.get = [mem](Env<T>& e, double* out) {
            for (size_t i = 0; i < (e.*mem).size(); i++) out[i] = (double)(e.*mem)[i];
        };
        fields[name] = f;
    }
    template <class S> void reg_scalar(const char* name, S Env<T>::*mem) {
        Field f;
        f.is_int = std::is_same<S, int32_t>::value;
        f.rows = [](Env<T>&) { return (size_t)1; };
        f.set = [mem](Env<T>& e, const double* in) { e.*mem = (S)in[0]; };
        f.get = [mem](Env<T>& e, double* out) { out[0] = (double)(e.*mem); };
        fields[name] = f;
    }

    void register_fields() {
#define R(n) reg(#n, &Env<T>::n)
        R(qpos); R(vel); R(acc); R(ctrl_pos); R(ctrl_vel); R(ctrl_force); R(qacc_ws); R(ctrl_mode);
        R(sort_value); R(sort_i_g); R(sort_is_max); R(cache_normal);
        R(links_pos); R(links_quat); R(root_COM); R(i_pos); R(i_quat); R(cinr_inertial); R(cinr_pos); R(cinr_mass);
        R(cd_ang); R(cd_vel); R(xanchor); R(xaxis); R(dof_pos); R(cdof_ang); R(cdof_vel); R(cdofd_ang); R(cdofd_vel);
        R(geoms_pos); R(geoms_quat); R(aabb_min); R(aabb_max);
        R(crb_inertial); R(crb_pos); R(crb_mass); R(f_ang); R(f_vel); R(mass_mat); R(mass_L); R(mass_D_inv);
        R(qf_applied); R(qf_passive); R(qf_bias); R(force); R(qf_smooth); R(acc_smooth); R(qf_constraint);
        R(cdd_ang); R(cdd_vel); R(cfrc_ang); R(cfrc_vel); R(cacc_ang); R(cacc_lin);
        R(broad_pairs); R(contact_geom_a); R(contact_geom_b); R(contact_link_a); R(contact_link_b); R(contact_pair);
        R(contact_pos); R(contact_normal); R(contact_penetration); R(contact_friction); R(contact_sol_params);
        R(contact_force);
        R(jac); R(diag); R(aref); R(efc_D); R(efc_frictionloss); R(Jaref); R(jv); R(efc_force); R(active);
        R(prev_active); R(qacc); R(Ma); R(grad); R(Mgrad); R(search); R(mv); R(qfrc_constraint); R(nt_H);
        R(cg_prev_grad); R(cg_prev_Mgrad);
#undef R
        reg_scalar("is_warmstart", &Env<T>::is_warmstart);
        reg_scalar("errno", &Env<T>::errno_);
        reg_scalar("first_time", &Env<T>::first_time);
        reg_scalar("n_broad", &Env<T>::n_broad);
        reg_scalar("n_contacts", &Env<T>::n_contacts);
        reg_scalar("n_gjk_regime", &Env<T>::n_gjk_regime);
        reg_scalar("n_gjk_regime_perturbed", &Env<T>::n_gjk_regime_perturbed);
        reg_scalar("n_con", &Env<T>::n_con);
        reg_scalar("n_con_eq", &Env<T>::n_con_eq);
        reg_scalar("n_con_fl", &Env<T>::n_con_fl);
        reg_scalar("improved", &Env<T>::improved);
        reg_scalar("n_iter", &Env<T>::n_iter);
        reg_scalar("cost", &Env<T>::cost);
        reg_scalar("cg_beta", &Env<T>::cg_beta);
        reg_scalar("cg_pg_dot_pMg", &Env<T>::cg_pg_dot_pMg);
    }

    void init_env(Env<T>& s) {
        s.allocate(m);
        for (int q = 0; q < m.Q; q++) s.qpos[q] = m.qpos0[q];
        for (int l = 0; l < m.L; l++) {
            Env<T>::s3(s.links_pos, l, m.vec3(m.link_pos, l));
            Env<T>::s4(s.links_quat, l, m.quat4(m.link_quat, l));
        }
        s.first_time = 1;
        s.is_warmstart = 0;
        forward_kinematics(m, s);
        com_links(m, s);
        update_geoms(m, s, true);
        forward_velocity(m, s);
    }

    // One substep (rigid_solver.py:912-980), with kinematics refreshed at the end like kernel_step_2.
    void step_env(Env<T>& s) {
        const bool implicit = m.h.integrator == RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST;
        compute_mass_matrix(m, s, implicit);
        factor_mass(m, s);
        torque_and_passive(m, s);
        update_acc(m, s, false);
        update_force(m, s);
        bias_force_and_qacc(m, s);
        if (!m.h.disable_constraint) add_equality_constraints(m, s);
        if (m.h.enable_collision) {
            update_aabbs(m, s);
            if (m.P > 0) {
                broadphase_sap(m, s);
                narrowphase(m, s);
            }
        }
        if (!m.h.disable_constraint) {
            add_inequality_constraints(m, s);
            constraint_solve(m, s);
        }
        update_acc(m, s, true);
        if (m.h.integrator != RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) implicit_damping(m, s);
        integrate(m, s);
        forward_kinematics(m, s);
        com_links(m, s);
        update_geoms(m, s, false);
        forward_velocity(m, s);
    }

    void stage(Env<T>& s, int st) {
        const bool implicit = m.h.integrator == RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST;
        switch (st) {
            case RS_STAGE_FK:
                forward_kinematics(m, s); com_links(m, s); update_geoms(m, s, false); forward_velocity(m, s);
                break;
            case RS_STAGE_CRBA: compute_mass_matrix(m, s, implicit); break;
            case RS_STAGE_FACTOR: factor_mass(m, s); break;
            case RS_STAGE_RNE:
                torque_and_passive(m, s); update_acc(m, s, false); update_force(m, s); bias_force_and_qacc(m, s);
                break;
            case RS_STAGE_AABB: update_aabbs(m, s); break;
            case RS_STAGE_BROADPHASE: if (m.P > 0) broadphase_sap(m, s); break;
            case RS_STAGE_NARROWPHASE: if (m.P > 0) narrowphase(m, s); break;
            case RS_STAGE_CONSTRAINTS:
                if (!m.h.disable_constraint) { add_equality_constraints(m, s); add_inequality_constraints(m, s); }
                break;
            case RS_STAGE_SOLVE: if (!m.h.disable_constraint) constraint_solve(m, s); break;
            case RS_STAGE_INTEGRATE:
                update_acc(m, s, true);
                if (m.h.integrator != RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) implicit_damping(m, s);
                integrate(m, s);
                forward_kinematics(m, s); com_links(m, s); update_geoms(m, s, false); forward_velocity(m, s);
                break;
        }
    }
};

struct Handle {
    int precision;
    void* ctx;
};

}  // namespace orc

using namespace orc;

template <class T> static int with_ctx(void* h, std::function<int(Context<T>&)> f) {
    return f(*reinterpret_cast<Context<T>*>(reinterpret_cast<Handle*>(h)->ctx));
}
// precision 64: double, 32: float, 0: the operation-counting double (oracle_flops.h)
#define DISPATCH(h, body)                                                              \
    (reinterpret_cast<Handle*>(h)->precision == 64                                     \
         ? with_ctx<double>(h, [&](Context<double>& c) -> int { body })                \
         : reinterpret_cast<Handle*>(h)->precision == 0                                \
               ? with_ctx<CF>(h, [&](Context<CF>& c) -> int { body })                  \
               : with_ctx<float>(h, [&](Context<float>& c) -> int { body }))

template <class T>
static int pair_contact(Context<T>& c, int ga, int gb, const double* pose_a, const double* pose_b, int mpr,
                        double* out) {
    V3<T> pa = v3(T(pose_a[0]), T(pose_a[1]), T(pose_a[2]));
    Q4<T> qa{T(pose_a[3]), T(pose_a[4]), T(pose_a[5]), T(pose_a[6])};
    V3<T> pb = v3(T(pose_b[0]), T(pose_b[1]), T(pose_b[2]));
    Q4<T> qb{T(pose_b[3]), T(pose_b[4]), T(pose_b[5]), T(pose_b[6])};
    V3<T> n = v3(T(0), T(0), T(0)), p = n;
    T pen = T(0);
    bool col = false;
    if (mpr) {
        mpr_contact(c.m, ga, gb, n, pa, qa, pb, qb, col, n, pen, p);
    } else {
        static thread_local GJKState<T> gjk;
        col = gjk_contact(c.m, gjk, ga, gb, pa, qa, pb, qb, n, p, pen);
    }
    for (int k = 0; k < 3; k++) out[k] = double(n[k]), out[3 + k] = double(p[k]);
    out[6] = double(pen);
    return col ? 1 : 0;
}

extern "C" {

void* orc_create(const rs_model_header* hdr, const void* blob, int precision, int n_envs) {
    Handle* h = new Handle{precision, nullptr};
    if (precision == 64) {
        auto* c = new Context<double>();
        if (!load_model(c->m, hdr, (const uint8_t*)blob)) return nullptr;
        c->register_fields();
        c->envs.resize(n_envs);
        for (auto& e : c->envs) c->init_env(e);
        h->ctx = c;
    } else if (precision == 0) {
        auto* c = new Context<CF>();
        if (!load_model(c->m, hdr, (const uint8_t*)blob)) return nullptr;
        c->register_fields();
        c->envs.resize(n_envs);
        for (auto& e : c->envs) c->init_env(e);
        h->ctx = c;
    } else {
        auto* c = new Context<float>();
        if (!load_model(c->m, hdr, (const uint8_t*)blob)) return nullptr;
        c->register_fields();
        c->envs.resize(n_envs);
        for (auto& e : c->envs) c->init_env(e);
        h->ctx = c;
    }
    return h;
}

// opts = {dt, gx, gy, gz, tolerance, ls_tolerance, meaninertia, mc_perturbation, mc_tolerance}
int orc_set_options(void* h, const double* o) {
    return DISPATCH(h, {
        c.m.dt = o[0];
        for (int k = 0; k < 3; k++) c.m.gravity[k] = o[1 + k];
        c.m.tolerance = o[4];
        c.m.ls_tolerance = o[5];
        c.m.meaninertia = o[6];
        c.m.mc_perturbation = o[7];
        c.m.mc_tolerance = o[8];
        return 0;
    });
}

void orc_destroy(void* h) {
    Handle* hh = reinterpret_cast<Handle*>(h);
    if (hh->precision == 64) delete reinterpret_cast<Context<double>*>(hh->ctx);
    else if (hh->precision == 0) delete reinterpret_cast<Context<CF>*>(hh->ctx);
    else delete reinterpret_cast<Context<float>*>(hh->ctx);
    delete hh;
}

int orc_n_envs(void* h) { return DISPATCH(h, return (int)c.envs.size();); }

// rows of a field per environment, or -1
int orc_field_rows(void* h, const char* name) {
    return DISPATCH(h, {
        auto it = c.fields.find(name);
        if (it == c.fields.end()) return -1;
        return (int)it->second.rows(c.envs[0]);
    });
}

int orc_field_is_int(void* h, const char* name) {
    return DISPATCH(h, {
        auto it = c.fields.find(name);
        if (it == c.fields.end()) return -1;
        return it->second.is_int;
    });
}

// data layout: [n_envs][rows]
int orc_get(void* h, const char* name, double* out) {
    return DISPATCH(h, {
        auto it = c.fields.find(name);
        if (it == c.fields.end()) return -1;
        size_t rows = it->second.rows(c.envs[0]);
        for (size_t b = 0; b < c.envs.size(); b++) it->second.get(c.envs[b], out + b * rows);
        return 0;
    });
}

int orc_set(void* h, const char* name, const double* in) {
    return DISPATCH(h, {
        auto it = c.fields.find(name);
        if (it == c.fields.end()) return -1;
        size_t rows = it->second.rows(c.envs[0]);
        for (size_t b = 0; b < c.envs.size(); b++) it->second.set(c.envs[b], in + b * rows);
        return 0;
    });
}

// Recompute kinematics (FK, COM, all geoms, velocities) of every environment from its qpos/vel.
int orc_forward_kinematics(void* h) {
    return DISPATCH(h, {
        for (auto& s : c.envs) {
            forward_kinematics(c.m, s); com_links(c.m, s); update_geoms(c.m, s, true); forward_velocity(c.m, s);
        }
        return 0;
    });
}

int orc_stage(void* h, int stage) {
    return DISPATCH(h, {
        for (auto& s : c.envs) c.stage(s, stage);
        return 0;
    });
}

// n_steps substeps on every environment using n_threads host threads (environments are independent).
int orc_step(void* h, int n_steps, int n_threads) {
    return DISPATCH(h, {
        int nt = n_threads < 1 ? 1 : n_threads;
        auto work = [&](int t) {
            for (size_t b = t; b < c.envs.size(); b += nt)
                for (int k = 0; k < n_steps; k++) c.step_env(c.envs[b]);
        };
        if (nt == 1) {
            work(0);
        } else {
            std::vector<std::thread> th;
            for (int t = 0; t < nt; t++) th.emplace_back(work, t);
            for (auto& x : th) x.join();
        }
        return 0;
    });
}

// One detection of geoms ga, gb of the model at the given world poses (pose = x, y, z, qw, qx, qy, qz) with the
// GJK/EPA path (gjk.py:170-446). out = normal[3], position[3], penetration; returns 1 when they collide. With
// mpr = 1 the MPR path (mpr.py:769-818) runs instead, without a warm-start direction.
int orc_pair_contact(void* h, int ga, int gb, const double* pose_a, const double* pose_b, int mpr, double* out) {
    if (reinterpret_cast<Handle*>(h)->precision == 0) return -1;
    if (reinterpret_cast<Handle*>(h)->precision == 64)
        return pair_contact<double>(*reinterpret_cast<Context<double>*>(reinterpret_cast<Handle*>(h)->ctx), ga, gb,
                                    pose_a, pose_b, mpr, out);
    return pair_contact<float>(*reinterpret_cast<Context<float>*>(reinterpret_cast<Handle*>(h)->ctx), ga, gb, pose_a,
                               pose_b, mpr, out);
}

// Operations counted so far by precision-0 handles (oracle_flops.h): out[0] all, out[1] library calls.
void orc_flops(uint64_t* out) {
    out[0] = FlopCounter::ops;
    out[1] = FlopCounter::transcendental;
}

}  // extern "C"
