// ORACLE (test infrastructure only): collision detection of one environment.
// Restates genesis/engine/solvers/rigid/collider/{broadphase,narrowphase,mpr,support_field,contact,box_contact,
// capsule_contact,utils}.py on the CPU code path (SAP broadphase + monolithic narrowphase, collider.py:684-808),
// including the GJK/EPA hand-over of the MPR path (narrowphase.py:726-851, oracle_gjk.h).
#pragma once
#include "oracle_gjk.h"
#include "oracle_model.h"

namespace orc {

template <class T> struct CCD {
    static T eps() { return sizeof(T) == 4 ? T(1e-9) : T(1e-10); }  // mpr.py:17
    static constexpr double tol = 1e-6;
    static constexpr int iters = 50;
};

// utils.py func_is_geom_aabbs_overlap
template <class T> bool aabb_overlap(const Env<T>& s, int a, int b) {
    for (int k = 0; k < 3; k++) {
        if (s.aabb_max[3 * a + k] <= s.aabb_min[3 * b + k]) return false;
        if (s.aabb_min[3 * a + k] >= s.aabb_max[3 * b + k]) return false;
    }
    return true;
}

// broadphase.py:74-138 func_collision_clear + :141-396 _func_broad_phase_sap (no hibernation)
template <class T> void broadphase_sap(const Model<T>& m, Env<T>& s) {
    for (int i = 0; i < s.n_contacts; i++) {
        s.contact_link_a[i] = s.contact_link_b[i] = s.contact_geom_a[i] = s.contact_geom_b[i] = -1;
        s.contact_penetration[i] = 0;
        for (int k = 0; k < 3; k++) s.contact_pos[3 * i + k] = s.contact_normal[3 * i + k] = s.contact_force[3 * i + k] = 0;
    }
    s.n_contacts = 0;
    s.n_gjk_regime = 0;
    s.n_gjk_regime_perturbed = 0;
    const int G = m.G;
    if (s.first_time) {
        int ib = 0;
        for (int l = 0; l < m.L; l++)
            for (int g = m.link_geom_start[l]; g < m.link_geom_end[l]; g++) {
                s.sort_value[2 * ib] = s.aabb_min[3 * g];
                s.sort_i_g[2 * ib] = g;
                s.sort_is_max[2 * ib] = 0;
                s.sort_value[2 * ib + 1] = s.aabb_max[3 * g];
                s.sort_i_g[2 * ib + 1] = g;
                s.sort_is_max[2 * ib + 1] = 1;
                ib++;
            }
        s.first_time = 0;
    } else {
        for (int i = 0; i < 2 * G; i++) {
            int g = s.sort_i_g[i];
            s.sort_value[i] = s.sort_is_max[i] ? s.aabb_max[3 * g] : s.aabb_min[3 * g];
        }
    }
    for (int i = 1; i < 2 * G; i++) {
        T kv = s.sort_value[i];
        int km = s.sort_is_max[i], kg = s.sort_i_g[i];
        int j = i - 1;
        while (j >= 0 && kv < s.sort_value[j]) {
            s.sort_value[j + 1] = s.sort_value[j];
            s.sort_is_max[j + 1] = s.sort_is_max[j];
            s.sort_i_g[j + 1] = s.sort_i_g[j];
            j--;
        }
        s.sort_value[j + 1] = kv;
        s.sort_is_max[j + 1] = km;
        s.sort_i_g[j + 1] = kg;
    }
    int n_broad = 0, n_active = 0;
    for (int i = 0; i < 2 * G; i++) {
        if (!s.sort_is_max[i]) {
            for (int j = 0; j < n_active; j++) {
                int ga = s.active_buffer[j], gb = s.sort_i_g[i];
                if (ga > gb) std::swap(ga, gb);
                int pidx = m.collision_pair_idx[ga * G + gb];
                if (pidx == -1) continue;
                if (!aabb_overlap(s, ga, gb)) {
                    for (int k = 0; k < 3; k++) s.cache_normal[3 * pidx + k] = T(0);
                    continue;
                }
                if (n_broad == m.BROAD) {
                    s.errno_ |= RS_ERR_OVERFLOW_CANDIDATE_CONTACTS;
                    break;
                }
                s.broad_pairs[2 * n_broad] = ga;
                s.broad_pairs[2 * n_broad + 1] = gb;
                n_broad++;
            }
            s.active_buffer[n_active++] = s.sort_i_g[i];
        } else {
            int rem = s.sort_i_g[i];
            for (int j = 0; j < n_active; j++)
                if (s.active_buffer[j] == rem) {
                    for (int k = j; k < n_active - 1; k++) s.active_buffer[k] = s.active_buffer[k + 1];
                    n_active--;
                    break;
                }
        }
    }
    s.n_broad = n_broad;
}

// ------------------------------------------------------------------------------------------ support functions

// support_field.py:139-181 _func_support_mesh (table lookup, 4 neighbouring cells)
template <class T> V3<T> support_mesh(const Model<T>& m, V3<T> d, int g) {
    const int res = 180;
    const T pi = T(3.141592653589793);
    T theta = std::atan2(d[1], d[0]);
    T phi = std::acos(std::fmin(std::fmax(d[2], T(-1)), T(1)));
    T ii = (theta + pi) / pi / T(2) * T(res);
    T jj = phi / pi * T(res);
    T dot_max = T(-1e20);
    V3<T> v = v3(T(0), T(0), T(0));
    for (int i4 = 0; i4 < 4; i4++) {
        int i, j;
        if (i4 % 2) i = int(std::fmod(std::ceil(ii), T(res)));
        else i = int(std::fmod(std::floor(ii), T(res)));
        if (i4 / 2 > 0) {
            j = int(std::fmin(std::fmax(std::ceil(jj), T(0)), T(res - 1)));
            if (j == res - 1) j = res - 2;
        } else {
            j = int(std::fmin(std::fmax(std::floor(jj), T(0)), T(res - 1)));
            if (j == 0) j = 1;
        }
        if (!(i >= 0 && i < res)) i = 0;  // NaN direction: keep the lookup inside the table
        int idx = m.geom_support_start[g] + i * res + j;
        V3<T> p = v3(m.support_v[3 * idx], m.support_v[3 * idx + 1], m.support_v[3 * idx + 2]);
        T dt = dot(p, d);
        if (dt > dot_max) {
            v = p;
            dot_max = dt;
        }
    }
    return v;
}

// mpr.py:147-202 support_driver
template <class T> V3<T> support_driver(const Model<T>& m, V3<T> d, int g, V3<T> pos, Q4<T> quat) {
    int t = m.geom_type[g];
    const T* data = &m.geom_data[7 * g];
    if (t == GT_SPHERE) return pos + d * data[0];
    if (t == GT_CAPSULE) {
        V3<T> axis = transform_by_quat(v3(T(0), T(0), T(1)), quat);
        T side = dot(d, axis) < T(0) ? T(-1) : T(1);
        V3<T> ep = pos + (T(0.5) * data[1]) * side * axis;
        return ep + d * data[0];
    }
    if (t == GT_BOX) {
        V3<T> db = inv_transform_by_quat(d, quat);
        V3<T> vl = v3((db[0] < T(0) ? T(-1) : T(1)) * data[0] * T(0.5), (db[1] < T(0) ? T(-1) : T(1)) * data[1] * T(0.5),
                      (db[2] < T(0) ? T(-1) : T(1)) * data[2] * T(0.5));
        return transform_by_trans_quat(vl, pos, quat);
    }
    if (m.geom_support_start[g] < 0) return pos;
    V3<T> dm = transform_by_quat(d, inv_quat(quat));
    return transform_by_trans_quat(support_mesh(m, dm, g), pos, quat);
}

// ------------------------------------------------------------------------------------------------------ MPR
template <class T> struct MPRState {
    V3<T> v[4], v1[4], v2[4];
    int size;
};

template <class T>
void compute_support(const Model<T>& m, V3<T> d, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb, V3<T>& v,
                     V3<T>& v1, V3<T>& v2) {
    v1 = support_driver(m, d, ga, pa, qa);
    v2 = support_driver(m, -d, gb, pb, qb);
    v = v1 - v2;
}

template <class T> V3<T> portal_dir(const MPRState<T>& st) {
    return normalized(cross(st.v[2] - st.v[1], st.v[3] - st.v[1]));
}

template <class T> bool reach_tolerance(const MPRState<T>& st, V3<T> v, V3<T> dir) {
    T dv1 = dot(st.v[1], dir), dv2 = dot(st.v[2], dir), dv3 = dot(st.v[3], dir), dv4 = dot(v, dir);
    T dot1 = std::fmin(std::fmin(dv4 - dv1, dv4 - dv2), dv4 - dv3);
    return dot1 < T(CCD<T>::tol) + CCD<T>::eps() * std::fmax(T(1), dot1);
}

template <class T> void expand_portal(MPRState<T>& st, V3<T> v, V3<T> v1, V3<T> v2) {
    V3<T> v4v0 = cross(v, st.v[0]);
    T d = dot(st.v[1], v4v0);
    int is;
    if (d > 0) {
        d = dot(st.v[2], v4v0);
        is = d > 0 ? 1 : 3;
    } else {
        d = dot(st.v[3], v4v0);
        is = d > 0 ? 2 : 1;
    }
    st.v1[is] = v1;
    st.v2[is] = v2;
    st.v[is] = v;
}

// mpr.py:446-599 mpr_discover_portal
template <class T>
int discover_portal(const Model<T>& m, MPRState<T>& st, int ga, int gb, V3<T> ca, V3<T> cb, V3<T> pa, Q4<T> qa,
                    V3<T> pb, Q4<T> qb) {
    const T E = CCD<T>::eps();
    st.v1[0] = ca;
    st.v2[0] = cb;
    st.v[0] = ca - cb;
    st.size = 1;
    if (std::fabs(st.v[0][0]) < E && std::fabs(st.v[0][1]) < E && std::fabs(st.v[0][2]) < E) st.v[0][0] += T(10) * E;
    V3<T> dir = -normalized(st.v[0]);
    V3<T> v, v1, v2;
    compute_support(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
    st.v1[1] = v1; st.v2[1] = v2; st.v[1] = v;
    st.size = 2;
    T d = dot(v, dir);
    if (d < E) return -1;
    dir = cross(st.v[0], st.v[1]);
    if (dot(dir, dir) < E) {
        if (std::fabs(st.v[1][0]) < E && std::fabs(st.v[1][1]) < E && std::fabs(st.v[1][2]) < E) return 1;
        return 2;
    }
    dir = normalized(dir);
    compute_support(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
    d = dot(v, dir);
    if (d < E) return -1;
    st.v1[2] = v1; st.v2[2] = v2; st.v[2] = v;
    st.size = 3;
    dir = normalized(cross(st.v[1] - st.v[0], st.v[2] - st.v[0]));
    d = dot(dir, st.v[0]);
    if (d > 0) {
        std::swap(st.v1[1], st.v1[2]);
        std::swap(st.v2[1], st.v2[2]);
        std::swap(st.v[1], st.v[2]);
        dir = -dir;
    }
    int trials = 0;
    while (st.size < 4) {
        compute_support(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
        d = dot(v, dir);
        if (d < E) return -1;
        bool cont = false;
        if (dot(cross(st.v[1], v), st.v[0]) < -E) {
            st.v1[2] = v1; st.v2[2] = v2; st.v[2] = v;
            cont = true;
        }
        if (!cont && dot(cross(v, st.v[2]), st.v[0]) < -E) {
            st.v1[1] = v1; st.v2[1] = v2; st.v[1] = v;
            cont = true;
        }
        if (cont) {
            dir = normalized(cross(st.v[1] - st.v[0], st.v[2] - st.v[0]));
            trials++;
            if (trials == 15) return -1;
        } else {
            st.v1[3] = v1; st.v2[3] = v2; st.v[3] = v;
            st.size = 4;
        }
    }
    return 0;
}

// mpr.py:233-276 mpr_refine_portal
template <class T>
int refine_portal(const Model<T>& m, MPRState<T>& st, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb) {
    const T E = CCD<T>::eps();
    // The reference loops until a return; the cap only bounds pathological inputs (the device kernel uses the same).
    for (int it = 0; it < 1024; it++) {
        V3<T> dir = portal_dir(st);
        if (dot(st.v[1], dir) > -E) return 0;
        V3<T> v, v1, v2;
        compute_support(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
        if (!(dot(v, dir) > -E) || reach_tolerance(st, v, dir)) return -1;
        expand_portal(st, v, v1, v2);
    }
    return -1;
}

// mpr.py:279-311 mpr_find_pos (non-MuJoCo-compatible branch)
template <class T> V3<T> find_pos(const MPRState<T>& st) {
    T b[4] = {0, 0, 0, 0};
    T sum = b[0] + b[1] + b[2] + b[3];
    if (sum < CCD<T>::eps()) {
        V3<T> dir = portal_dir(st);
        b[0] = 0;
        for (int i = 1; i < 4; i++) {
            int i1 = i % 3 + 1, i2 = (i + 1) % 3 + 1;
            b[i] = dot(cross(st.v[i1], st.v[i2]), dir);
        }
        sum = b[0] + b[1] + b[2] + b[3];
    }
    V3<T> p1 = v3(T(0), T(0), T(0)), p2 = v3(T(0), T(0), T(0));
    for (int i = 0; i < 4; i++) {
        p1 = p1 + b[i] * st.v1[i];
        p2 = p2 + b[i] * st.v2[i];
    }
    return (T(0.5) / sum) * (p1 + p2);
}

// mpr.py:339-421 mpr_find_penetration
template <class T>
void find_penetration(const Model<T>& m, MPRState<T>& st, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb,
                      bool& is_col, V3<T>& normal, T& pen, V3<T>& pos) {
    int it = 0;
    while (true) {
        V3<T> dir = portal_dir(st);
        V3<T> v, v1, v2;
        compute_support(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
        if (reach_tolerance(st, v, dir) || it > CCD<T>::iters) {
            pen = dot(dir, st.v[1]);
            normal = -dir;
            is_col = true;
            pos = find_pos(st);
            return;
        }
        expand_portal(st, v, v1, v2);
        it++;
    }
}

// mpr.py:602-702 guess_geoms_center
template <class T>
void guess_centers(const Model<T>& m, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb, V3<T> nws, V3<T>& ca,
                   V3<T>& cb) {
    ca = transform_by_trans_quat(m.vec3(m.geom_center, ga), pa, qa);
    cb = transform_by_trans_quat(m.vec3(m.geom_center, gb), pb, qb);
    const T E = CCD<T>::eps();
    if (std::fabs(nws[0]) > E || std::fabs(nws[1]) > E || std::fabs(nws[2]) > E) {
        const T* A = &m.geom_init_aabb[24 * ga];
        const T* B = &m.geom_init_aabb[24 * gb];
        V3<T> a0 = v3(A[0], A[1], A[2]), a7 = v3(A[21], A[22], A[23]);
        V3<T> b0 = v3(B[0], B[1], B[2]), b7 = v3(B[21], B[22], B[23]);
        ca = transform_by_trans_quat(T(0.5) * (a7 + a0), pa, qa);
        cb = transform_by_trans_quat(T(0.5) * (b7 + b0), pb, qb);
        V3<T> delta = ca - cb;
        V3<T> nrm = normalized(delta);
        if (norm(cross(nws, nrm)) > T(0.01)) {
            V3<T> off = dot(delta, nws) * nws - delta;
            T on = norm(off);
            if (on > m.EPS) {
                V3<T> dir = off / on;
                V3<T> dla = inv_transform_by_quat(dir, qa), dlb = inv_transform_by_quat(dir, qb);
                V3<T> sa = a7 - a0, sb = b7 - b0;
                T la = dot(sa, v3(std::fabs(dla[0]), std::fabs(dla[1]), std::fabs(dla[2])));
                T lb = dot(sb, v3(std::fabs(dlb[0]), std::fabs(dlb[1]), std::fabs(dlb[2])));
                T ratio = std::fmin(on / (la + lb), T(0.5));
                ca = ca + dir * la * ratio;
                cb = cb - dir * lb * ratio;
            }
        }
    }
}

// mpr.py:705-818 func_mpr_contact_from_centers / func_mpr_contact
template <class T>
void mpr_contact(const Model<T>& m, int ga, int gb, V3<T> nws, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb, bool& is_col,
                 V3<T>& normal, T& pen, V3<T>& pos) {
    V3<T> ca, cb;
    guess_centers(m, ga, gb, pa, qa, pb, qb, nws, ca, cb);
    MPRState<T> st;
    int res = discover_portal(m, st, ga, gb, ca, cb, pa, qa, pb, qb);
    is_col = false;
    pos = normal = v3(T(0), T(0), T(0));
    pen = 0;
    if (res == 1) {
        is_col = true;
        pen = 0;
        normal = -normalized(st.v[0]);
        pos = (st.v1[1] + st.v2[1]) * T(0.5);
    } else if (res == 2) {
        is_col = true;
        pen = norm(st.v[1]);
        normal = -normalized(st.v[1]);
        pos = (st.v1[1] + st.v2[1]) * T(0.5);
    } else if (res == 0) {
        res = refine_portal(m, st, ga, gb, pa, qa, pb, qb);
        if (res >= 0) find_penetration(m, st, ga, gb, pa, qa, pb, qb, is_col, normal, pen, pos);
    }
}

// ------------------------------------------------------------------------------------------ contacts helpers

// contact.py:318-336 func_compute_tolerance
template <class T> T compute_tolerance(const Model<T>& m, int ga, int gb, T tol) {
    auto sz = [&](int g) {
        const T* A = &m.geom_init_aabb[24 * g];
        return norm(v3(A[21] - A[0], A[22] - A[1], A[23] - A[2]));
    };
    T s = sz(gb);
    if (m.geom_type[ga] != GT_PLANE) s = std::fmin(sz(ga), s);
    return T(0.5) * tol * s;
}

// contact.py:181-222 func_add_contact
template <class T> void add_contact(const Model<T>& m, Env<T>& s, int ga, int gb, V3<T> n, V3<T> p, T pen, int pair) {
    int i = s.n_contacts;
    if (i < m.KMAX) {
        T fa = m.geom_friction[ga] * T(1), fb = m.geom_friction[gb] * T(1);
        s.contact_geom_a[i] = ga;
        s.contact_geom_b[i] = gb;
        Env<T>::s3(s.contact_normal, i, n);
        Env<T>::s3(s.contact_pos, i, p);
        s.contact_penetration[i] = pen;
        s.contact_friction[i] = std::fmax(std::fmax(fa, fb), T(1e-2));
        for (int k = 0; k < 7; k++)
            s.contact_sol_params[7 * i + k] = T(0.5) * (m.geom_sol_params[7 * ga + k] + m.geom_sol_params[7 * gb + k]);
        s.contact_link_a[i] = m.geom_link[ga];
        s.contact_link_b[i] = m.geom_link[gb];
        s.contact_pair[i] = pair;
        s.n_contacts = i + 1;
    } else {
        s.errno_ |= RS_ERR_OVERFLOW_COLLISION_PAIRS;
    }
}

// contact.py:354-412 func_contact_orthogonals (non-MuJoCo-compatible branch)
template <class T>
void contact_orthogonals(const Model<T>& m, const Env<T>& s, int ga, int gb, V3<T> normal, V3<T>& ax0, V3<T>& ax1) {
    int g = gb;
    if (m.geom_type[ga] != GT_PLANE) {
        const T* A = &m.geom_init_aabb[24 * ga + 21];
        const T* B = &m.geom_init_aabb[24 * gb + 21];
        T va = A[0] * A[1] * A[2], vb = B[0] * B[1] * B[2];
        g = va < vb ? ga : gb;
    }
    int l = m.geom_link[g];
    M3<T> R = quat_to_R(Env<T>::g4(s.i_quat, l), m.EPS);
    int idx = 0;
    T amax = 0;
    for (int i = 0; i < 3; i++) {
        T a = std::fabs(dot(v3(R.m[0][i], R.m[1][i], R.m[2][i]), normal));
        if (a > amax) {
            amax = a;
            idx = i;
        }
    }
    idx = (idx + 1) % 3;
    ax0 = v3(R.m[0][idx], R.m[1][idx], R.m[2][idx]);
    ax0 = normalized(ax0 - dot(normal, ax0) * normal);
    ax1 = cross(normal, ax0);
}

// contact.py:416-436 func_rotate_frame
template <class T> void rotate_frame(V3<T> pos, Q4<T> quat, V3<T> cpos, Q4<T> qrot, V3<T>& npos, Q4<T>& nquat) {
    nquat = transform_quat_by_quat(quat, qrot);
    V3<T> rel = cpos - pos;
    V3<T> vec = transform_by_quat(rel, qrot) - rel;
    npos = pos - vec;
}

// utils.py:9-68 func_closest_points_on_segments
template <class T> void closest_points_segments(V3<T> a1, V3<T> a2, V3<T> b1, V3<T> b2, T EPS, V3<T>& pa, V3<T>& pb) {
    V3<T> da = a2 - a1, db = b2 - b1, r = a1 - b1;
    T aa = dot(da, da), ab = dot(da, db), bb = dot(db, db), d = dot(da, r), e = dot(db, r);
    T denom = aa * bb - ab * ab;
    T sp = 0, t = 0;
    if (denom < EPS) {
        sp = 0;
        t = bb > EPS ? clampv(e / bb, T(0), T(1)) : T(0);
    } else {
        sp = (ab * e - bb * d) / denom;
        t = (aa * e - ab * d) / denom;
        sp = clampv(sp, T(0), T(1));
        t = clampv(bb > EPS ? (ab * sp + e) / bb : T(0), T(0), T(1));
        T sn = clampv(aa > EPS ? (ab * t - d) / aa : T(0), T(0), T(1));
        if (aa > EPS) sp = sn;
    }
    pa = a1 + sp * da;
    pb = b1 + t * db;
}

// geom.py:34-44 qd_transform_by_quat_fast
template <class T> V3<T> transform_by_quat_fast(V3<T> v, Q4<T> q) {
    V3<T> u = v3(q.x, q.y, q.z);
    V3<T> t = T(2) * cross(u, v);
    return v + q.w * t + cross(u, t);
}

// capsule_contact.py:10-99
template <class T>
bool capsule_capsule(const Model<T>& m, int ga, int gb, V3<T> pa_, Q4<T> qa, V3<T> pb_, Q4<T> qb, V3<T>& n, V3<T>& cp,
                     T& pen) {
    T ra = m.geom_data[7 * ga], ha = T(0.5) * m.geom_data[7 * ga + 1];
    T rb = m.geom_data[7 * gb], hb = T(0.5) * m.geom_data[7 * gb + 1];
    V3<T> z = v3(T(0), T(0), T(1));
    V3<T> axa = transform_by_quat_fast(z, qa), axb = transform_by_quat_fast(z, qb);
    V3<T> P1 = pa_ - ha * axa, P2 = pa_ + ha * axa, Q1 = pb_ - hb * axb, Q2 = pb_ + hb * axb;
    V3<T> Pa, Pb;
    closest_points_segments(P1, P2, Q1, Q2, m.EPS, Pa, Pb);
    V3<T> diff = Pa - Pb;
    T d2 = dot(diff, diff);
    T cr = ra + rb;
    n = v3(T(1), T(0), T(0));
    cp = v3(T(0), T(0), T(0));
    pen = 0;
    if (!(d2 < cr * cr)) return false;
    T dist = std::sqrt(d2);
    if (dist > m.EPS) {
        n = diff / dist;
    } else {
        V3<T> nd = cross(axa, axb);
        T ndl = dot(nd, nd);
        if (ndl > m.EPS) n = nd / ndl;
        else n = std::fabs(axa[0]) < T(0.9) ? v3(T(0), -axa[2], axa[1]) : v3(axa[2], T(0), -axa[0]);
    }
    pen = cr - dist;
    cp = Pa - ra * n + T(0.5) * pen * n;
    return true;
}

// capsule_contact.py:102-196
template <class T>
bool sphere_capsule(const Model<T>& m, int ga, int gb, V3<T> pa_, Q4<T> qa, V3<T> pb_, Q4<T> qb, V3<T>& n, V3<T>& cp,
                    T& pen) {
    T ndir = 1;
    V3<T> sc = pa_, cc = pb_;
    Q4<T> cq = qb;
    if (m.geom_type[gb] == GT_SPHERE) {
        std::swap(ga, gb);
        sc = pb_;
        cc = pa_;
        cq = qa;
        ndir = -1;
    }
    T sr = m.geom_data[7 * ga];
    T cr = m.geom_data[7 * gb], ch = T(0.5) * m.geom_data[7 * gb + 1];
    V3<T> ax = transform_by_quat_fast(v3(T(0), T(0), T(1)), cq);
    V3<T> P1 = cc - ch * ax, P2 = cc + ch * ax;
    V3<T> seg = P2 - P1;
    T sl = dot(seg, seg);
    T t = T(0.5);
    if (sl > m.EPS) t = clampv(dot(sc - P1, seg) / sl, T(0), T(1));
    V3<T> cl = P1 + t * seg;
    V3<T> diff = sc - cl;
    T d2 = dot(diff, diff);
    T comb = sr + cr;
    n = v3(T(1), T(0), T(0));
    cp = v3(T(0), T(0), T(0));
    pen = 0;
    bool col = false;
    if (d2 < comb * comb) {
        col = true;
        T dist = std::sqrt(d2);
        if (dist > m.EPS) n = diff / dist;
        else n = std::fabs(ax[0]) < T(0.9) ? v3(T(0), -ax[2], ax[1]) : v3(ax[2], T(0), -ax[0]);
        pen = comb - dist;
        cp = sc - (sr - T(0.5) * pen) * n;
    }
    n = n * ndir;
    return col;
}

// narrowphase.py:520-970 func_convex_convex_contact (MPR with the GJK/EPA hand-over)
template <class T> void convex_convex_contact(const Model<T>& m, Env<T>& s, int ga, int gb) {
    using E = Env<T>;
    int ta = m.geom_type[ga], tb = m.geom_type[gb];
    if (ta == GT_PLANE && tb == GT_BOX) return;
    const T EPS = m.EPS;
    const T mc_pert = T(m.mc_perturbation), mc_tol = T(m.mc_tolerance);
    bool multi = m.h.enable_multi_contact && ta != GT_SPHERE && ta != GT_ELLIPSOID && tb != GT_SPHERE &&
                 tb != GT_ELLIPSOID;
    T tolerance = compute_tolerance(m, ga, gb, mc_tol);
    V3<T> pa0 = E::g3(s.geoms_pos, ga), pb0 = E::g3(s.geoms_pos, gb);
    Q4<T> qa0 = E::g4(s.geoms_quat, ga), qb0 = E::g4(s.geoms_quat, gb);
    V3<T> pa = pa0, pb = pb0;
    Q4<T> qa = qa0, qb = qb0;
    bool is_col0 = false, is_col = false;
    T pen0 = 0, pen = 0;
    V3<T> n0 = v3(T(0), T(0), T(0)), p0 = n0, n = n0, cp = n0, ax0 = n0, ax1 = n0;
    Q4<T> qrot{T(0), T(0), T(0), T(0)};
    int n_con = 0;
    int pair = m.collision_pair_idx[(ga < gb ? ga : gb) * m.G + (ga < gb ? gb : ga)];
    for (int det = 0; det < 5; det++) {
        if (multi && is_col0) {
            V3<T> axis = T(2 * (det % 2) - 1) * ax0 + T(1 - 2 * ((det / 2) % 2)) * ax1;
            qrot = rotvec_to_quat(mc_pert * axis, EPS);
            rotate_frame(pa0, qa0, p0, qrot, pa, qa);
            rotate_frame(pb0, qb0, p0, inv_quat(qrot), pb, qb);
        }
        if ((multi && is_col0) || det == 0) {
            if (ta == GT_CAPSULE && tb == GT_CAPSULE) {
                is_col = capsule_capsule(m, ga, gb, pa, qa, pb, qb, n, cp, pen);
            } else if ((ta == GT_SPHERE && tb == GT_CAPSULE) || (ta == GT_CAPSULE && tb == GT_SPHERE)) {
                is_col = sphere_capsule(m, ga, gb, pa, qa, pb, qb, n, cp, pen);
            } else if (ta == GT_PLANE) {
                V3<T> pd = transform_by_quat(v3(m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]), qa);
                n = -normalized(pd);
                V3<T> v1 = support_driver(m, n, gb, pb, qb);
                pen = dot(n, v1 - pa);
                cp = v1 - T(0.5) * pen * n;
                is_col = pen > T(0);
            } else {
                bool updated = false;
                V3<T> nws = E::g3(s.cache_normal, pair);
                bool avail = std::fabs(nws[0]) > EPS || std::fabs(nws[1]) > EPS || std::fabs(nws[2]) > EPS;
                for (int i_mpr = 0; i_mpr < 2; i_mpr++) {
                    if (i_mpr == 1 && det == 0 && !is_col && avail) {
                        nws = v3(T(0), T(0), T(0));
                        avail = false;
                        updated = false;
                    }
                    if (!updated) {
                        mpr_contact(m, ga, gb, nws, pa, qa, pb, qb, is_col, n, pen, cp);
                        updated = true;
                    }
                }
                // narrowphase.py:726-734: a deep penetration with no cached direction, or a large overlap, is
                // detected again with GJK/EPA (narrowphase.py:766-851; a GJK detection gives one contact, whose
                // normal, position and depth replace MPR's when the geoms collide)
                if (pen > tolerance &&
                    (!avail || T(m.h.mc_tolerance) * pen >= T(m.h.mpr_to_gjk_overlap_ratio) * tolerance)) {
                    if (det == 0) s.n_gjk_regime++;
                    else s.n_gjk_regime_perturbed++;
                    static thread_local GJKState<T> gjk;
                    V3<T> gn, gp;
                    T gpen;
                    is_col = gjk_contact(m, gjk, ga, gb, pa, qa, pb, qb, gn, gp, gpen);
                    pen = gpen;
                    if (is_col) {
                        n = gn;
                        cp = gp;
                    }
                }
            }
        }
        if (det == 0) {
            is_col0 = is_col;
            n0 = n;
            pen0 = pen;
            p0 = cp;
            if (is_col0) {
                add_contact(m, s, ga, gb, n0, p0, pen0, pair);
                if (multi) {
                    contact_orthogonals(m, s, ga, gb, n, ax0, ax1);
                    n_con = 1;
                }
                E::s3(s.cache_normal, pair, n);
            } else {
                E::s3(s.cache_normal, pair, v3(T(0), T(0), T(0)));
            }
        } else if (multi && is_col) {
            V3<T> cpa = transform_by_quat((cp - T(0.5) * pen * n) - p0, inv_quat(qrot)) + p0;
            V3<T> cpb = transform_by_quat((cp + T(0.5) * pen * n) - p0, qrot) + p0;
            cp = T(0.5) * (cpa + cpb);
            V3<T> tw = cross(n, n0);
            for (int k = 0; k < 3; k++) tw[k] = clampv(tw[k], -mc_pert, mc_pert);
            n = n + cross(tw, n);
            pen = dot(n, cpb - cpa);
            bool repeated = false;
            for (int i = 0; i < n_con; i++) {
                if (!repeated) {
                    int ip = s.n_contacts - 1 - i;
                    if (ip >= 0 && norm(cp - E::g3(s.contact_pos, ip)) < tolerance) repeated = true;
                }
            }
            if (!repeated && pen > -tolerance) {
                pen = std::fmax(pen, T(0));
                add_contact(m, s, ga, gb, n, cp, pen, pair);
                n_con++;
            }
        }
    }
}

// box_contact.py:25-97 func_plane_box_contact
template <class T> void plane_box_contact(const Model<T>& m, Env<T>& s, int ga, int gb) {
    using E = Env<T>;
    V3<T> pa = E::g3(s.geoms_pos, ga), pb = E::g3(s.geoms_pos, gb);
    Q4<T> qa = E::g4(s.geoms_quat, ga), qb = E::g4(s.geoms_quat, gb);
    int pair = m.collision_pair_idx[(ga < gb ? ga : gb) * m.G + (ga < gb ? gb : ga)];
    V3<T> pd = transform_by_quat(v3(m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]), qa);
    V3<T> n = -normalized(pd);
    V3<T> v1 = support_driver(m, n, gb, pb, qb);
    T pen = dot(n, v1 - pa);
    if (!(pen > T(0))) return;
    V3<T> cp = v1 - T(0.5) * pen * n;
    add_contact(m, s, ga, gb, n, cp, pen, pair);
    if (!m.h.enable_multi_contact) return;
    int n_con = 1;
    V3<T> cp0 = cp;
    T tol = compute_tolerance(m, ga, gb, T(m.mc_tolerance));
    const T* d = &m.geom_data[7 * gb];
    for (int c = 0; c < 8; c++) {
        if (n_con >= m.h.n_contacts_per_pair) break;
        // trimesh box vertex order: x slowest, z fastest
        V3<T> loc = v3((c & 4 ? T(0.5) : T(-0.5)) * d[0], (c & 2 ? T(0.5) : T(-0.5)) * d[1], (c & 1 ? T(0.5) : T(-0.5)) * d[2]);
        V3<T> corner = transform_by_trans_quat(loc, pb, qb);
        T p = dot(n, corner - pa);
        if (p > T(0)) {
            V3<T> c2 = corner - T(0.5) * p * n;
            if (norm(c2 - cp0) > tol) {
                add_contact(m, s, ga, gb, n, c2, p, pair);
                n_con++;
            }
        }
    }
}

// contact.py:16-31 rotaxis / rotmatx
template <class T> inline V3<T> rotaxis(V3<T> v, int i0, int i1, int i2, T f0, T f1, T f2) {
    return v3(v[i0] * f0, v[i1] * f1, v[i2] * f2);
}
template <class T> inline M3<T> rotmatx(const M3<T>& a, int i0, int i1, int i2, T f0, T f1, T f2) {
    M3<T> r;
    for (int c = 0; c < 3; c++) {
        r.m[0][c] = a.m[i0][c] * f0;
        r.m[1][c] = a.m[i1][c] * f1;
        r.m[2][c] = a.m[i2][c] * f2;
    }
    return r;
}
template <class T> inline M3<T> mtrans(const M3<T>& a) {
    M3<T> r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[j][i];
    return r;
}
template <class T> inline M3<T> mmul(const M3<T>& a, const M3<T>& b) {
    M3<T> r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j];
    return r;
}
template <class T> inline V3<T> mvec(const M3<T>& a, V3<T> v) {
    return v3(a.m[0][0] * v[0] + a.m[0][1] * v[1] + a.m[0][2] * v[2], a.m[1][0] * v[0] + a.m[1][1] * v[1] + a.m[1][2] * v[2],
              a.m[2][0] * v[0] + a.m[2][1] * v[1] + a.m[2][2] * v[2]);
}
template <class T> inline V3<T> mrow(const M3<T>& a, int i) { return v3(a.m[i][0], a.m[i][1], a.m[i][2]); }

// box_contact.py:251-303 / :488-520: the frame that puts the clipping face on +z (face index q in 0..5)
template <class T> inline void box_face_frame(int q, M3<T>& rotmore, int& i0, int& i1, int& i2, T& f0, T& f1, T& f2) {
    rotmore = M3<T>{{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}}};
    i0 = 0, i1 = 1, i2 = 2;
    f0 = f1 = f2 = T(1);
    if (q == 0) rotmore.m[0][2] = T(-1), rotmore.m[1][1] = T(1), rotmore.m[2][0] = T(1), i0 = 2, f0 = T(-1), i2 = 0;
    else if (q == 1) rotmore.m[0][0] = T(1), rotmore.m[1][2] = T(-1), rotmore.m[2][1] = T(1), i1 = 2, f1 = T(-1), i2 = 1;
    else if (q == 2) rotmore.m[0][0] = T(1), rotmore.m[1][1] = T(1), rotmore.m[2][2] = T(1);
    else if (q == 3) rotmore.m[0][2] = T(1), rotmore.m[1][1] = T(1), rotmore.m[2][0] = T(-1), i0 = 2, i2 = 0, f2 = T(-1);
    else if (q == 4) rotmore.m[0][0] = T(1), rotmore.m[1][2] = T(1), rotmore.m[2][1] = T(-1), i1 = 2, i2 = 1, f2 = T(-1);
    else if (q == 5) rotmore.m[0][0] = T(-1), rotmore.m[1][1] = T(1), rotmore.m[2][2] = T(-1), f0 = T(-1), f2 = T(-1);
}

// box_contact.py:100-785 func_box_box_contact (MuJoCo's engine_collision_box.c separating-axis test, then face or edge
// clipping). The reference keeps up to n_contacts_per_pair candidate points and writes past that array when a
// configuration yields more (it does not check); this restatement keeps every candidate (at most 24).
template <class T> void box_box_contact(const Model<T>& m, Env<T>& s, int ga, int gb) {
    using E = Env<T>;
    const T EPS = m.EPS;
    const T margin = T(0), margin2 = margin * margin;
    int pair = m.collision_pair_idx[(ga < gb ? ga : gb) * m.G + (ga < gb ? gb : ga)];
    int n = 0, code = -1, cle1 = 0, cle2 = 0;
    bool is_return = false, in_ = false;
    M3<T> rotmore{{{T(0), T(0), T(0)}, {T(0), T(0), T(0)}, {T(0), T(0), T(0)}}};
    const V3<T> pos1 = E::g3(s.geoms_pos, ga), pos2 = E::g3(s.geoms_pos, gb);
    const Q4<T> q1_ = E::g4(s.geoms_quat, ga), q2_ = E::g4(s.geoms_quat, gb);
    const V3<T> size1 = v3(m.geom_data[7 * ga] / T(2), m.geom_data[7 * ga + 1] / T(2), m.geom_data[7 * ga + 2] / T(2));
    const V3<T> size2 = v3(m.geom_data[7 * gb] / T(2), m.geom_data[7 * gb + 1] / T(2), m.geom_data[7 * gb + 2] / T(2));
    const M3<T> mat1 = quat_to_R(q1_, EPS), mat2 = quat_to_R(q2_, EPS);
    const V3<T> pos21 = mvec(mtrans(mat1), pos2 - pos1);
    const V3<T> pos12 = mvec(mtrans(mat2), pos1 - pos2);
    const M3<T> rot = mmul(mtrans(mat1), mat2), rott = mtrans(rot);
    M3<T> rotabs, rottabs;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) rotabs.m[i][j] = std::fabs(rot.m[i][j]), rottabs.m[i][j] = std::fabs(rott.m[i][j]);
    const V3<T> plen2 = mvec(rotabs, size2);
    const V3<T> plen1 = mvec(mtrans(rotabs), size1);
    T penetration = margin;
    for (int i = 0; i < 3; i++) penetration = penetration + size1[i] * T(3) + size2[i] * T(3);
    // face axes of either box (box_contact.py:162-175)
    for (int i = 0; i < 3; i++) {
        T c1 = -std::fabs(pos21[i]) + size1[i] + plen2[i];
        T c2 = -std::fabs(pos12[i]) + size2[i] + plen1[i];
        if (c1 < -margin || c2 < -margin) is_return = true;
        if (c1 < penetration) {
            penetration = c1;
            code = i + 3 * (pos21[i] < T(0) ? 1 : 0) + 0;
        }
        if (c2 < penetration) {
            penetration = c2;
            code = i + 3 * (pos12[i] < T(0) ? 1 : 0) + 6;
        }
    }
    // edge-edge axes (box_contact.py:176-238)
    V3<T> clnorm = v3(T(0), T(0), T(0));
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            const T rj0 = rott.m[j][0], rj1 = rott.m[j][1], rj2 = rott.m[j][2];
            V3<T> tmp2 = i == 0 ? v3(T(0), -rj2, rj1) : (i == 1 ? v3(rj2, T(0), -rj0) : v3(-rj1, rj0, T(0)));
            const T c1 = norm(tmp2);
            tmp2 = tmp2 / c1;
            if (c1 >= EPS) {
                const T c2 = dot(pos21, tmp2);
                T c3 = T(0);
                for (int k = 0; k < 3; k++)
                    if (k != i) c3 = c3 + size1[k] * std::fabs(tmp2[k]);
                for (int k = 0; k < 3; k++)
                    if (k != j) c3 = c3 + size2[k] * rotabs.m[i][3 - k - j] / c1;
                c3 = c3 - std::fabs(c2);
                if (c3 < -margin) is_return = true;
                if (c3 < penetration * (T(1) - T(1e-12))) {
                    penetration = c3;
                    cle1 = 0;
                    for (int k = 0; k < 3; k++)
                        if (k != i && ((tmp2[k] > T(0)) != (c2 < T(0)))) cle1 += 1 << k;
                    cle2 = 0;
                    for (int k = 0; k < 3; k++)
                        if (k != j && (((rot.m[i][3 - k - j] > T(0)) != (c2 < T(0))) != ((k - j + 3) % 3 == 1)))
                            cle2 += 1 << k;
                    code = 12 + i * 3 + j;
                    clnorm = tmp2;
                    in_ = c2 < T(0);
                }
            }
        }
    }
    if (code == -1) is_return = true;
    if (is_return) return;
    V3<T> points[24];
    T depth[24];
    if (code < 12) {
        // face contact (box_contact.py:241-457)
        const int q1 = code % 6, q2 = code / 6;
        int i0, i1, i2;
        T f0, f1, f2;
        box_face_frame(q1, rotmore, i0, i1, i2, f0, f1, f2);
        M3<T> r;
        V3<T> p, tmp1, sz;
        if (q2) {
            r = mmul(rotmore, mtrans(rot));
            p = rotaxis(pos12, i0, i1, i2, f0, f1, f2);
            tmp1 = rotaxis(size2, i0, i1, i2, f0, f1, f2);
            sz = size1;
        } else {
            r = rotmatx(rot, i0, i1, i2, f0, f1, f2);
            p = rotaxis(pos21, i0, i1, i2, f0, f1, f2);
            tmp1 = rotaxis(size1, i0, i1, i2, f0, f1, f2);
            sz = size2;
        }
        const M3<T> rt = mtrans(r);
        const V3<T> ss = v3(std::fabs(tmp1[0]), std::fabs(tmp1[1]), std::fabs(tmp1[2]));
        const T lx = ss[0], ly = ss[1], hz = ss[2];
        p[2] = p[2] - hz;
        V3<T> lp = p;
        int clcorner = 0;
        for (int i = 0; i < 3; i++)
            if (r.m[2][i] < T(0)) clcorner += 1 << i;
        for (int i = 0; i < 3; i++) lp = lp + mrow(rt, i) * sz[i] * T((clcorner & (1 << i)) ? 1 : -1);
        V3<T> pts[6];
        for (int i = 0; i < 6; i++) pts[i] = v3(T(0), T(0), T(0));
        int mm = 0, k = 0;
        pts[mm++] = lp;
        for (int i = 0; i < 3; i++)
            if (std::fabs(r.m[2][i]) < T(0.5)) pts[mm++] = mrow(rt, i) * sz[i] * T((clcorner & (1 << i)) ? -2 : 2);
        pts[3] = pts[0] + pts[1];
        pts[4] = pts[0] + pts[2];
        pts[5] = pts[3] + pts[2];
        V3<T> la[4], lb[4];
        if (mm > 1) la[k] = pts[0], lb[k] = pts[1], k++;
        if (mm > 2) {
            la[k] = pts[0], lb[k] = pts[2], k++;
            la[k] = pts[3], lb[k] = pts[2], k++;
            la[k] = pts[4], lb[k] = pts[1], k++;
        }
        for (int i = 0; i < k; i++)
            for (int q = 0; q < 2; q++) {
                const T a = la[i][q], b = lb[i][q], c = la[i][1 - q], d = lb[i][1 - q];
                if (std::fabs(b) > EPS)
                    for (int jj = 0; jj < 2; jj++) {
                        const T l = ss[q] * T(2 * jj - 1);
                        const T c1 = (l - a) / b;
                        if (T(0) <= c1 && c1 <= T(1)) {
                            const T c2 = c + d * c1;
                            if (std::fabs(c2) <= ss[1 - q]) points[n++] = la[i] + lb[i] * c1;
                        }
                    }
            }
        const T a = pts[1][0], b = pts[2][0], c = pts[1][1], d = pts[2][1];
        const T c1 = a * d - b * c;
        if (mm > 2)
            for (int i = 0; i < 4; i++) {
                const T llx = (i / 2) ? lx : -lx, lly = (i % 2) ? ly : -ly;
                const T x = llx - pts[0][0], y = lly - pts[0][1];
                const T u = (x * d - y * b) / c1, v = (y * a - x * c) / c1;
                if (T(0) < u && u < T(1) && T(0) < v && v < T(1))
                    points[n++] = v3(llx, lly, pts[0][2] + u * pts[1][2] + v * pts[2][2]);
            }
        for (int i = 0; i < (1 << (mm - 1)); i++) {
            const V3<T> t = pts[i == 0 ? 0 : i + 2];
            if (!(i && (t[0] <= -lx || t[0] >= lx || t[1] <= -ly || t[1] >= ly))) points[n++] = t;
        }
        mm = n;
        n = 0;
        for (int i = 0; i < mm; i++)
            if (points[i][2] <= margin) {
                points[n] = points[i];
                depth[n] = points[n][2];
                points[n][2] = points[n][2] * T(0.5);
                n++;
            }
        const M3<T> rr = mmul(q2 ? mat2 : mat1, mtrans(rotmore));
        const V3<T> pp = q2 ? pos2 : pos1;
        const T sg = q2 ? T(-1) : T(1);
        const V3<T> normal0 = v3(sg * rr.m[0][2], sg * rr.m[1][2], sg * rr.m[2][2]);
        int n_added = 0;
        const int n_start = s.n_contacts;
        for (int i = 0; i < n; i++) {
            if (n_added < m.h.n_contacts_per_pair) {
                // the reference takes the halved height as the distance here (box_contact.py:434)
                const T dist = points[i][2];
                points[i][2] = points[i][2] + hz;
                const V3<T> cpos = pp + mvec(rr, points[i]);
                bool valid = true;
                for (int j = 0; j < n_added && n_start + j < m.KMAX; j++) {
                    const V3<T> q = E::g3(s.contact_pos, n_start + j);
                    if (std::fabs(cpos[0] - q[0]) < EPS && std::fabs(cpos[1] - q[1]) < EPS && std::fabs(cpos[2] - q[2]) < EPS)
                        valid = false;
                }
                if (valid) {
                    add_contact(m, s, ga, gb, -normal0, cpos, -dist, pair);
                    n_added++;
                }
            }
        }
        return;
    }
    // edge-edge contact (box_contact.py:458-785)
    code -= 12;
    const int q1 = code / 3, q2 = code % 3;
    int ax1 = 0, ax2 = 0, pax1 = 0, pax2 = 0;
    if (q2 == 0) ax1 = 1, ax2 = 2;
    else if (q2 == 1) ax1 = 0, ax2 = 2;
    else ax1 = 1, ax2 = 0;
    if (q1 == 0) pax1 = 1, pax2 = 2;
    else if (q1 == 1) pax1 = 0, pax2 = 2;
    else pax1 = 1, pax2 = 0;
    if (rotabs.m[q1][ax1] < rotabs.m[q1][ax2]) {
        ax1 = ax2;
        ax2 = 3 - q2 - ax1;
    }
    if (rottabs.m[q2][pax1] < rottabs.m[q2][pax2]) {
        pax1 = pax2;
        pax2 = 3 - q1 - pax1;
    }
    const int clface = (cle1 & (1 << pax2)) ? pax2 : pax2 + 3;
    int i0, i1, i2;
    T f0, f1, f2;
    box_face_frame(clface, rotmore, i0, i1, i2, f0, f1, f2);
    V3<T> p = rotaxis(pos21, i0, i1, i2, f0, f1, f2);
    const V3<T> rnorm = rotaxis(clnorm, i0, i1, i2, f0, f1, f2);
    const M3<T> r = rotmatx(rot, i0, i1, i2, f0, f1, f2);
    const V3<T> tmp1s = mvec(mtrans(rotmore), size1);
    const V3<T> sz = v3(std::fabs(tmp1s[0]), std::fabs(tmp1s[1]), std::fabs(tmp1s[2]));
    const M3<T> rt = mtrans(r);
    const T lx = sz[0], ly = sz[1], hz = sz[2];
    p[2] = p[2] - hz;
    const T e1 = (cle2 & (1 << ax1)) ? T(1) : T(-1), e2 = (cle2 & (1 << ax2)) ? T(1) : T(-1);
    V3<T> pq[4];
    pq[0] = p;
    pq[0] = pq[0] + mrow(rt, ax1) * size2[ax1] * e1;
    pq[0] = pq[0] + mrow(rt, ax2) * size2[ax2] * e2;
    pq[1] = pq[0];
    pq[0] = pq[0] + mrow(rt, q2) * size2[q2];
    pq[1] = pq[1] - mrow(rt, q2) * size2[q2];
    pq[2] = p;
    pq[2] = pq[2] + mrow(rt, ax1) * size2[ax1] * (-e1);
    pq[2] = pq[2] + mrow(rt, ax2) * size2[ax2] * e2;
    pq[3] = pq[2];
    pq[2] = pq[2] + mrow(rt, q2) * size2[q2];
    pq[3] = pq[3] - mrow(rt, q2) * size2[q2];
    const V3<T> axi0 = pq[0], axi1 = pq[1] - pq[0], axi2 = pq[2] - pq[0];
    if (std::fabs(rnorm[2]) < EPS) return;
    const T innorm = (T(1) / rnorm[2]) * (in_ ? T(-1) : T(1));
    V3<T> pu[4];
    T ppts2[4][2];
    for (int i = 0; i < 4; i++) {
        const T c1 = -pq[i][2] / rnorm[2];
        pu[i] = pq[i];
        pq[i] = pq[i] + c1 * rnorm;
        ppts2[i][0] = pq[i][0];
        ppts2[i][1] = pq[i][1];
    }
    const V3<T> pts0 = pq[0], pts1 = pq[1] - pq[0], pts2 = pq[2] - pq[0];
    V3<T> la[4] = {pts0, pts0, pts0 + pts1, pts0 + pts2}, lb[4] = {pts1, pts2, pts2, pts1};
    V3<T> ua[4] = {axi0, axi0, axi0 + axi1, axi0 + axi2}, ub[4] = {axi1, axi2, axi2, axi1};
    for (int i = 0; i < 4; i++)
        for (int q = 0; q < 2; q++) {
            const T a = la[i][q], b = lb[i][q], c = la[i][1 - q], d = lb[i][1 - q];
            if (std::fabs(b) > EPS)
                for (int jj = 0; jj < 2; jj++) {
                    const T l = sz[q] * T(2 * jj - 1);
                    const T c1 = (l - a) / b;
                    if (T(0) <= c1 && c1 <= T(1)) {
                        const T c2 = c + d * c1;
                        if (std::fabs(c2) <= sz[1 - q] && (ua[i][2] + ub[i][2] * c1) * innorm <= margin) {
                            points[n] = ua[i] * T(0.5) + c1 * T(0.5) * ub[i];
                            points[n][q] = points[n][q] + T(0.5) * l;
                            points[n][1 - q] = points[n][1 - q] + T(0.5) * c2;
                            depth[n] = points[n][2] * innorm * T(2);
                            n++;
                        }
                    }
                }
        }
    const int nl = n;
    const T a = pts1[0], b = pts2[0], c = pts1[1], d = pts2[1];
    const T c1 = a * d - b * c;
    for (int i = 0; i < 4; i++) {
        const T llx = (i / 2) ? lx : -lx, lly = (i % 2) ? ly : -ly;
        const T x = llx - pts0[0], y = lly - pts0[1];
        T u = (x * d - y * b) / c1, v = (y * a - x * c) / c1;
        if (nl == 0) {
            if ((u < T(0) || u > T(1)) && (v < T(0) || v > T(1))) continue;
        } else if (u < T(0) || u > T(1) || v < T(0) || v > T(1)) {
            continue;
        }
        u = std::fmin(std::fmax(u, T(0)), T(1));
        v = std::fmin(std::fmax(v, T(0)), T(1));
        const V3<T> t1 = pu[0] * (T(1) - u - v) + pu[1] * u + pu[2] * v;
        points[n] = v3(llx, lly, T(0));
        const V3<T> t2 = points[n] - t1;
        const T cc = dot(t2, t2);
        if (!(t1[2] > T(0) && cc > margin2)) {
            points[n] = points[n] + t1;
            points[n] = points[n] * T(0.5);
            depth[n] = std::sqrt(cc) * (t1[2] < T(0) ? T(-1) : T(1));
            n++;
        }
    }
    const int nf = n;
    for (int i = 0; i < 4; i++) {
        const T x = ppts2[i][0], y = ppts2[i][1];
        if (nl == 0) {
            if (nf != 0 && (x < -lx || x > lx) && (y < -ly || y > ly)) continue;
        } else if (x < -lx || x > lx || y < -ly || y > ly) {
            continue;
        }
        T cc = T(0);
        for (int j = 0; j < 2; j++) {
            if (ppts2[i][j] < -sz[j]) cc = cc + (ppts2[i][j] + sz[j]) * (ppts2[i][j] + sz[j]);
            else if (ppts2[i][j] > sz[j]) cc = cc + (ppts2[i][j] - sz[j]) * (ppts2[i][j] - sz[j]);
        }
        cc = cc + (pu[i][2] * innorm) * (pu[i][2] * innorm);
        if (pu[i][2] > T(0) && cc > margin2) continue;
        V3<T> t1 = v3(ppts2[i][0] * T(0.5), ppts2[i][1] * T(0.5), T(0));
        for (int j = 0; j < 2; j++) {
            if (ppts2[i][j] < -sz[j]) t1[j] = -sz[j] * T(0.5);
            else if (ppts2[i][j] > sz[j]) t1[j] = sz[j] * T(0.5);
        }
        t1 = t1 + pu[i] * T(0.5);
        points[n] = t1;
        depth[n] = std::sqrt(cc) * (pu[i][2] < T(0) ? T(-1) : T(1));
        n++;
    }
    const M3<T> rr = mmul(mat1, mtrans(rotmore));
    const V3<T> nr = mvec(rr, rnorm);
    const V3<T> normal0 = (in_ ? T(-1) : T(1)) * nr;
    int n_added = 0;
    const int n_start = s.n_contacts;
    for (int i = 0; i < n; i++) {
        if (n_added < m.h.n_contacts_per_pair) {
            const T dist = depth[i];
            points[i][2] = points[i][2] + hz;
            const V3<T> cpos = pos1 + mvec(rr, points[i]);
            bool valid = true;
            for (int j = 0; j < n_added && n_start + j < m.KMAX; j++) {
                const V3<T> q = E::g3(s.contact_pos, n_start + j);
                if (std::fabs(cpos[0] - q[0]) < EPS && std::fabs(cpos[1] - q[1]) < EPS && std::fabs(cpos[2] - q[2]) < EPS)
                    valid = false;
            }
            if (valid) {
                add_contact(m, s, ga, gb, -normal0, cpos, -dist, pair);
                n_added++;
            }
        }
    }
}

// narrowphase.py:1988-2060 func_narrow_phase_convex_vs_convex + :2135-2189 specializations
template <class T> void narrowphase(const Model<T>& m, Env<T>& s) {
    for (int i = 0; i < s.n_broad; i++) {
        int ga = s.broad_pairs[2 * i], gb = s.broad_pairs[2 * i + 1];
        if (m.geom_type[ga] > m.geom_type[gb]) std::swap(ga, gb);
        int ta = m.geom_type[ga], tb = m.geom_type[gb];
        if (m.geom_is_convex[ga] && m.geom_is_convex[gb] && tb != GT_TERRAIN &&
            !(m.h.box_box_detection && ta == GT_BOX && tb == GT_BOX))
            if (!(ta == GT_PLANE && tb == GT_BOX)) convex_convex_contact(m, s, ga, gb);
    }
    for (int i = 0; i < s.n_broad; i++) {
        int ga = s.broad_pairs[2 * i], gb = s.broad_pairs[2 * i + 1];
        if (m.geom_type[ga] > m.geom_type[gb]) std::swap(ga, gb);
        if (m.geom_type[ga] == GT_PLANE && m.geom_type[gb] == GT_BOX) plane_box_contact(m, s, ga, gb);
        if (m.h.box_box_detection && m.geom_type[ga] == GT_BOX && m.geom_type[gb] == GT_BOX) box_box_contact(m, s, ga, gb);
    }
}

}  // namespace orc
