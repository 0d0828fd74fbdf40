// Collision detection of one environment (one lane per environment).
// Restates genesis/engine/solvers/rigid/collider/{broadphase,narrowphase,mpr,support_field,contact,box_contact,
// capsule_contact,utils}.py: warm-started sweep-and-prune, then the monolithic convex narrowphase (MPR with cached
// normal, 4 perturbed re-detections for multi-contact), then the plane-box specialisation. Contacts are appended in the
// broadphase pair order, which makes the contact list deterministic. Deep or poorly warm-started MPR detections are
// handed to GJK/EPA (narrowphase.py:726-851, rs_gjk.h).
#pragma once
#include "rs_device.h"
#include "rs_gjk.h"

namespace rs {

constexpr float CCD_EPS = 1e-9f;  // mpr.py:17 (FP32)
constexpr float CCD_TOL = 1e-6f;
constexpr int CCD_ITERS = 50;

template <class S>
__device__ inline bool aabb_overlap(const S& st, int b, int ga, int gb) {
    for (int k = 0; k < 3; k++) {
        if (SF(aabb_max, 3 * ga + k) <= SF(aabb_min, 3 * gb + k)) return false;
        if (SF(aabb_min, 3 * ga + k) >= SF(aabb_max, 3 * gb + k)) return false;
    }
    return true;
}

// broadphase.py:74-138 func_collision_clear: reset the contact rows [from, to) of the previous substep
template <class S>
__device__ inline void clear_contacts(const S& st, int b, int from, int to) {
    for (int i = from; i < to; i++) {
        SF(contact_link_a, i) = SF(contact_link_b, i) = SF(contact_geom_a, i) = SF(contact_geom_b, i) = -1;
        SF(contact_penetration, i) = 0.f;
        ST3(contact_pos, i, zero3());
        ST3(contact_normal, i, zero3());
        ST3(contact_force, i, zero3());
    }
}

// broadphase.py:141-200: first call fills the endpoint buffer in link/geom order; later calls refresh the values of
// the previous order (warm start)
template <class S>
__device__ inline void sap_init_endpoints(const DModel& m, const S& st, int b) {
    int ib = 0;
    for (int l = 0; l < m.L; l++)
        for (int g = m.link_geom_start[l]; g < m.link_geom_end[l]; g++) {
            SF(sort_value, 2 * ib) = SF(aabb_min, 3 * g);
            SF(sort_i_g, 2 * ib) = g;
            SF(sort_is_max, 2 * ib) = 0;
            SF(sort_value, 2 * ib + 1) = SF(aabb_max, 3 * g);
            SF(sort_i_g, 2 * ib + 1) = g;
            SF(sort_is_max, 2 * ib + 1) = 1;
            ib++;
        }
    SF(first_time, 0) = 0;
}

// broadphase.py:200-230: insertion sort of the endpoints (stable: equal values keep their previous order)
template <class S>
__device__ inline void sap_insertion_sort(const S& st, int b, int n) {
    for (int i = 1; i < n; i++) {
        float kv = SF(sort_value, i);
        int km = SF(sort_is_max, i), kg = SF(sort_i_g, i);
        int j = i - 1;
        while (j >= 0 && kv < SF(sort_value, j)) {
            SF(sort_value, j + 1) = SF(sort_value, j);
            SF(sort_is_max, j + 1) = SF(sort_is_max, j);
            SF(sort_i_g, j + 1) = SF(sort_i_g, j);
            j--;
        }
        SF(sort_value, j + 1) = kv;
        SF(sort_is_max, j + 1) = km;
        SF(sort_i_g, j + 1) = kg;
    }
}

// broadphase.py:141-396 _func_broad_phase_sap. (st, b) holds the AABBs, sort buffers, active list and pair list;
// (gs, ge) holds the warm-start normal cache (the same view in the per-stage kernels).
template <class S>
__device__ inline void broadphase_sap(const DModel& m, const S& st, int b, const DState& gs, int ge) {
    const int G = m.G;
    if (SF(first_time, 0)) {
        sap_init_endpoints(m, st, b);
    } else {
        for (int i = 0; i < 2 * G; i++) {
            int g = SF(sort_i_g, i);
            SF(sort_value, i) = SF(sort_is_max, i) ? SF(aabb_max, 3 * g) : SF(aabb_min, 3 * g);
        }
    }
    sap_insertion_sort(st, b, 2 * G);
    int n_broad = 0, n_active = 0;
    for (int i = 0; i < 2 * G; i++) {
        if (!SF(sort_is_max, i)) {
            int gi = SF(sort_i_g, i);
            for (int j = 0; j < n_active; j++) {
                int ga = SF(active_buffer, j), gb = gi;
                if (ga > gb) {
                    int t = ga;
                    ga = gb;
                    gb = t;
                }
                int pidx = m.collision_pair_idx[ga * G + gb];
                if (pidx == -1) continue;
                if (!aabb_overlap(st, b, ga, gb)) {
                    st3(gs.cache_normal, pidx, gs.B, ge, zero3());
                    continue;
                }
                if (n_broad == m.BROAD) {
                    SF(errbits, 0) = SF(errbits, 0) | RS_ERR_OVERFLOW_CANDIDATE_CONTACTS;
                    break;
                }
                SF(broad_pairs, 2 * n_broad) = ga;
                SF(broad_pairs, 2 * n_broad + 1) = gb;
                n_broad++;
            }
            SF(active_buffer, n_active) = gi;
            n_active++;
        } else {
            int rem = SF(sort_i_g, i);
            for (int j = 0; j < n_active; j++)
                if (SF(active_buffer, j) == rem) {
                    for (int k = j; k < n_active - 1; k++) SF(active_buffer, k) = SF(active_buffer, k + 1);
                    n_active--;
                    break;
                }
        }
    }
    SF(n_broad, 0) = n_broad;
}

template <class S>
__device__ inline void dev_broadphase(const DModel& m, const S& st, int b) {
    clear_contacts(st, b, 0, SF(n_contacts, 0));
    SF(n_contacts, 0) = 0;
    broadphase_sap(m, st, b, st, b);
}

// ----------------------------------------------------------------------------------------------- support

// support_field.py:139-181: 4-neighbour lookup in the 180x180 (theta, phi) table of a mesh geom
__device__ inline f3 support_mesh(const DModel& m, f3 d, int g) {
    const int res = 180;
    const float pi = 3.14159265358979323846f;
    float theta = atan2f(d.y, d.x);
    float phi = acosf(clampf(d.z, -1.f, 1.f));
    float ii = (theta + pi) / pi / 2.f * (float)res;
    float jj = phi / pi * (float)res;
    float dmax = -1e20f;
    f3 v = zero3();
    const int start = m.geom_support_start[g];
    const int vstart = m.geom_vert_start[g];
    for (int i4 = 0; i4 < 4; i4++) {
        // the reference's cell % res; theta lies in [-pi, pi], so the cell lies in [0, res] and the modulo is one
        // conditional subtraction (a NaN cell stays NaN, as with fmodf)
        float cell = (i4 & 1) ? ceilf(ii) : floorf(ii);
        if (cell >= (float)res) cell = cell - (float)res;
        int i = (int)cell;
        int j;
        if (i4 >> 1) {
            j = (int)clampf(ceilf(jj), 0.f, (float)(res - 1));
            if (j == res - 1) j = res - 2;
        } else {
            j = (int)clampf(floorf(jj), 0.f, (float)(res - 1));
            if (j == 0) j = 1;
        }
        // NaN directions map to cell (0, 1); keep the index inside the table in every case
        i = min(max(i, 0), res - 1);
        // the table cell holds the id of its vertex; the vertex itself comes from the (small) vertex array, so the
        // lookup touches 4 bytes of the large table instead of 12
        const f3 p = mvec3(m.verts, vstart + m.support_vid[(size_t)(start + i * res + j)]);
        float dt = dot3(p, d);
        if (dt > dmax) {
            v = p;
            dmax = dt;
        }
    }
    return v;
}

// mpr.py:147-202 support_driver
__device__ inline f3 support_driver(const DModel& m, f3 d, int g, f3 pos, f4q quat) {
    int t = m.geom_type[g];
    cfptr data = &m.geom_data[7 * g];
    if (t == GT_SPHERE) return pos + d * data[0];
    if (t == GT_CAPSULE) {
        f3 axis = qrot(mk3(0.f, 0.f, 1.f), quat);
        float side = dot3(d, axis) < 0.f ? -1.f : 1.f;
        f3 ep = pos + (0.5f * data[1]) * side * axis;
        return ep + d * data[0];
    }
    if (t == GT_BOX) {
        f3 db = qrot_inv(d, quat);
        f3 vl = mk3((db.x < 0.f ? -1.f : 1.f) * data[0] * 0.5f, (db.y < 0.f ? -1.f : 1.f) * data[1] * 0.5f,
                    (db.z < 0.f ? -1.f : 1.f) * data[2] * 0.5f);
        return tr_point(vl, pos, quat);
    }
    if (m.geom_support_start[g] < 0) return pos;  // no support table (not a mesh): never read outside the tables
    f3 dm = qrot(d, qinv(quat));
    return tr_point(support_mesh(m, dm, g), pos, quat);
}

// --------------------------------------------------------------------------------------------------- MPR
struct MPR {
    f3 v[4], v1[4], v2[4];
    int size;
};

__device__ inline void support_pair(const DModel& m, f3 d, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb, f3& v, f3& v1,
                                    f3& v2) {
    v1 = support_driver(m, d, ga, pa, qa);
    v2 = support_driver(m, -d, gb, pb, qb);
    v = v1 - v2;
}
__device__ inline f3 portal_dir(const MPR& s) { return normalize3(cross3(s.v[2] - s.v[1], s.v[3] - s.v[1])); }
__device__ inline bool reach_tol(const MPR& s, f3 v, f3 dir) {
    float d1 = dot3(s.v[1], dir), d2 = dot3(s.v[2], dir), d3 = dot3(s.v[3], dir), d4 = dot3(v, dir);
    float dt = fminf(fminf(d4 - d1, d4 - d2), d4 - d3);
    return dt < CCD_TOL + CCD_EPS * fmaxf(1.f, dt);
}
__device__ inline void expand_portal(MPR& s, f3 v, f3 v1, f3 v2) {
    f3 v4v0 = cross3(v, s.v[0]);
    float d = dot3(s.v[1], v4v0);
    int is;
    if (d > 0.f) {
        d = dot3(s.v[2], v4v0);
        is = d > 0.f ? 1 : 3;
    } else {
        d = dot3(s.v[3], v4v0);
        is = d > 0.f ? 2 : 1;
    }
    s.v1[is] = v1;
    s.v2[is] = v2;
    s.v[is] = v;
}

// mpr.py:446-599
__device__ inline int discover_portal(const DModel& m, MPR& s, int ga, int gb, f3 ca, f3 cb, f3 pa, f4q qa, f3 pb,
                                      f4q qb) {
    s.v1[0] = ca;
    s.v2[0] = cb;
    s.v[0] = ca - cb;
    s.size = 1;
    if (fabsf(s.v[0].x) < CCD_EPS && fabsf(s.v[0].y) < CCD_EPS && fabsf(s.v[0].z) < CCD_EPS) s.v[0].x += 10.f * CCD_EPS;
    f3 dir = -normalize3(s.v[0]);
    f3 v, v1, v2;
    support_pair(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
    s.v1[1] = v1;
    s.v2[1] = v2;
    s.v[1] = v;
    s.size = 2;
    if (dot3(v, dir) < CCD_EPS) return -1;
    dir = cross3(s.v[0], s.v[1]);
    if (dot3(dir, dir) < CCD_EPS) {
        if (fabsf(s.v[1].x) < CCD_EPS && fabsf(s.v[1].y) < CCD_EPS && fabsf(s.v[1].z) < CCD_EPS) return 1;
        return 2;
    }
    dir = normalize3(dir);
    support_pair(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
    if (dot3(v, dir) < CCD_EPS) return -1;
    s.v1[2] = v1;
    s.v2[2] = v2;
    s.v[2] = v;
    s.size = 3;
    dir = normalize3(cross3(s.v[1] - s.v[0], s.v[2] - s.v[0]));
    if (dot3(dir, s.v[0]) > 0.f) {
        f3 t;
        t = s.v1[1]; s.v1[1] = s.v1[2]; s.v1[2] = t;
        t = s.v2[1]; s.v2[1] = s.v2[2]; s.v2[2] = t;
        t = s.v[1]; s.v[1] = s.v[2]; s.v[2] = t;
        dir = -dir;
    }
    int trials = 0;
    while (s.size < 4) {
        support_pair(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
        if (dot3(v, dir) < CCD_EPS) return -1;
        bool cont = false;
        if (dot3(cross3(s.v[1], v), s.v[0]) < -CCD_EPS) {
            s.v1[2] = v1;
            s.v2[2] = v2;
            s.v[2] = v;
            cont = true;
        }
        if (!cont && dot3(cross3(v, s.v[2]), s.v[0]) < -CCD_EPS) {
            s.v1[1] = v1;
            s.v2[1] = v2;
            s.v[1] = v;
            cont = true;
        }
        if (cont) {
            dir = normalize3(cross3(s.v[1] - s.v[0], s.v[2] - s.v[0]));
            if (++trials == 15) return -1;
        } else {
            s.v1[3] = v1;
            s.v2[3] = v2;
            s.v[3] = v;
            s.size = 4;
        }
    }
    return 0;
}

// mpr.py:233-276; bounded by the same iteration cap as the penetration search so every lane terminates
__device__ inline int refine_portal(const DModel& m, MPR& s, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb) {
    for (int it = 0; it < 1024; it++) {
        f3 dir = portal_dir(s);
        if (dot3(s.v[1], dir) > -CCD_EPS) return 0;
        f3 v, v1, v2;
        support_pair(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
        if (!(dot3(v, dir) > -CCD_EPS) || reach_tol(s, v, dir)) return -1;
        expand_portal(s, v, v1, v2);
    }
    return -1;
}

__device__ inline f3 find_pos(const MPR& s) {
    float bw[4] = {0.f, 0.f, 0.f, 0.f};
    float sum = bw[0] + bw[1] + bw[2] + bw[3];
    if (sum < CCD_EPS) {
        f3 dir = portal_dir(s);
        bw[0] = 0.f;
        bw[1] = dot3(cross3(s.v[2], s.v[3]), dir);
        bw[2] = dot3(cross3(s.v[3], s.v[1]), dir);
        bw[3] = dot3(cross3(s.v[1], s.v[2]), dir);
        sum = bw[0] + bw[1] + bw[2] + bw[3];
    }
    f3 p1 = zero3(), p2 = zero3();
    for (int i = 0; i < 4; i++) {
        p1 = p1 + bw[i] * s.v1[i];
        p2 = p2 + bw[i] * s.v2[i];
    }
    return (0.5f / sum) * (p1 + p2);
}

__device__ inline void find_penetration(const DModel& m, MPR& s, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb,
                                        bool& is_col, f3& n, float& pen, f3& pos) {
    for (int it = 0;; it++) {
        f3 dir = portal_dir(s);
        f3 v, v1, v2;
        support_pair(m, dir, ga, gb, pa, qa, pb, qb, v, v1, v2);
        if (reach_tol(s, v, dir) || it > CCD_ITERS) {
            pen = dot3(dir, s.v[1]);
            n = -dir;
            is_col = true;
            pos = find_pos(s);
            return;
        }
        expand_portal(s, v, v1, v2);
    }
}

// mpr.py:602-702 guess_geoms_center
__device__ inline void guess_centers(const DModel& m, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb, f3 nws, f3& ca,
                                     f3& cb, float eps) {
    ca = tr_point(mvec3(m.geom_center, ga), pa, qa);
    cb = tr_point(mvec3(m.geom_center, gb), pb, qb);
    if (fabsf(nws.x) > CCD_EPS || fabsf(nws.y) > CCD_EPS || fabsf(nws.z) > CCD_EPS) {
        cfptr A = &m.geom_init_aabb[24 * ga];
        cfptr Bx = &m.geom_init_aabb[24 * gb];
        f3 a0 = mk3(A[0], A[1], A[2]), a7 = mk3(A[21], A[22], A[23]);
        f3 b0 = mk3(Bx[0], Bx[1], Bx[2]), b7 = mk3(Bx[21], Bx[22], Bx[23]);
        ca = tr_point(0.5f * (a7 + a0), pa, qa);
        cb = tr_point(0.5f * (b7 + b0), pb, qb);
        f3 delta = ca - cb;
        f3 nrm = normalize3(delta);
        if (norm3(cross3(nws, nrm)) > 0.01f) {
            f3 off = dot3(delta, nws) * nws - delta;
            float on = norm3(off);
            if (on > eps) {
                f3 dir = off / on;
                f3 dla = qrot_inv(dir, qa), dlb = qrot_inv(dir, qb);
                f3 sa = a7 - a0, sb = b7 - b0;
                float la = dot3(sa, mk3(fabsf(dla.x), fabsf(dla.y), fabsf(dla.z)));
                float lb = dot3(sb, mk3(fabsf(dlb.x), fabsf(dlb.y), fabsf(dlb.z)));
                float ratio = fminf(on / (la + lb), 0.5f);
                ca = ca + dir * la * ratio;
                cb = cb - dir * lb * ratio;
            }
        }
    }
}

__device__ inline void mpr_contact(const DModel& m, int ga, int gb, f3 nws, f3 pa, f4q qa, f3 pb, f4q qb, bool& is_col,
                                   f3& n, float& pen, f3& pos) {
    f3 ca, cb;
    guess_centers(m, ga, gb, pa, qa, pb, qb, nws, ca, cb, m.eps);
    MPR s;
    int res = discover_portal(m, s, ga, gb, ca, cb, pa, qa, pb, qb);
    is_col = false;
    n = pos = zero3();
    pen = 0.f;
    if (res == 1) {
        is_col = true;
        n = -normalize3(s.v[0]);
        pos = (s.v1[1] + s.v2[1]) * 0.5f;
    } else if (res == 2) {
        is_col = true;
        pen = norm3(s.v[1]);
        n = -normalize3(s.v[1]);
        pos = (s.v1[1] + s.v2[1]) * 0.5f;
    } else if (res == 0) {
        if (refine_portal(m, s, ga, gb, pa, qa, pb, qb) >= 0) find_penetration(m, s, ga, gb, pa, qa, pb, qb, is_col, n, pen, pos);
    }
}

// ------------------------------------------------------------------------------------------------- contacts
__device__ inline float geom_tolerance(const DModel& m, int ga, int gb, float tol) {
    cfptr B7 = &m.geom_init_aabb[24 * gb];
    float s = norm3(mk3(B7[21] - B7[0], B7[22] - B7[1], B7[23] - B7[2]));
    if (m.geom_type[ga] != GT_PLANE) {
        cfptr A = &m.geom_init_aabb[24 * ga];
        s = fminf(norm3(mk3(A[21] - A[0], A[22] - A[1], A[23] - A[2])), s);
    }
    return 0.5f * tol * s;
}

// contact.py:181-222 func_add_contact: write contact row i of (st, b)
template <class S>
__device__ inline void write_contact(const DModel& m, const S& st, int b, int i, int ga, int gb, f3 n, f3 p,
                                     float pen, int pair) {
    float fa = m.geom_friction[ga] * 1.f, fb = m.geom_friction[gb] * 1.f;
    SF(contact_geom_a, i) = ga;
    SF(contact_geom_b, i) = gb;
    ST3(contact_normal, i, n);
    ST3(contact_pos, i, p);
    SF(contact_penetration, i) = pen;
    SF(contact_friction, i) = fmaxf(fmaxf(fa, fb), 1e-2f);
    for (int k = 0; k < 7; k++)
        SF(contact_sol_params, 7 * i + k) = 0.5f * (m.geom_sol_params[7 * ga + k] + m.geom_sol_params[7 * gb + k]);
    SF(contact_link_a, i) = m.geom_link[ga];
    SF(contact_link_b, i) = m.geom_link[gb];
    SF(contact_pair, i) = pair;
}

template <class S>
__device__ inline void add_contact(const DModel& m, const S& st, int b, int ga, int gb, f3 n, f3 p, float pen,
                                   int pair) {
    int i = SF(n_contacts, 0);
    if (i < m.KMAX) {
        write_contact(m, st, b, i, ga, gb, n, p, pen, pair);
        SF(n_contacts, 0) = i + 1;
    } else {
        SF(errbits, 0) = SF(errbits, 0) | RS_ERR_OVERFLOW_COLLISION_PAIRS;
    }
}

// Contact sinks used by the pair routines. PlaneSink appends to the contact rows of one environment (the sequential
// func_add_contact); SlotSink keeps one pair's contacts in a private slot array (8 floats per contact: normal, pos,
// penetration) so that pairs can run on different lanes and be compacted in pair order afterwards.
struct PlaneSink {
    const DModel& m;
    const DState& st;
    int b;
    __device__ void add(int ga, int gb, f3 n, f3 p, float pen, int pair) { add_contact(m, st, b, ga, gb, n, p, pen, pair); }
    // true when contact (count - 1 - back) exists and lies within tol of p
    __device__ bool near(int back, f3 p, float tol) const {
        int ip = SF(n_contacts, 0) - 1 - back;
        return ip >= 0 && norm3(p - LD3(contact_pos, ip)) < tol;
    }
    __device__ int count() const { return SF(n_contacts, 0); }
    // true when contact i exists and every coordinate of it lies within eps of p
    __device__ bool same(int i, f3 p, float eps) const {
        if (i >= SF(n_contacts, 0)) return false;
        const f3 q = LD3(contact_pos, i);
        return fabsf(p.x - q.x) < eps && fabsf(p.y - q.y) < eps && fabsf(p.z - q.z) < eps;
    }
};

struct SlotSink {
    float* slot;  // capacity: n_contacts_per_pair entries of 8 floats
    int cap;
    int n;
    __device__ void add(int, int, f3 nrm, f3 p, float pen, int) {
        if (n < cap) {
            float* s = slot + 8 * n;
            s[0] = nrm.x; s[1] = nrm.y; s[2] = nrm.z;
            s[3] = p.x; s[4] = p.y; s[5] = p.z;
            s[6] = pen;
            n++;
        }
    }
    __device__ bool near(int back, f3 p, float tol) const {
        int ip = n - 1 - back;
        if (ip < 0) return false;
        const float* s = slot + 8 * ip;
        return norm3(p - mk3(s[3], s[4], s[5])) < tol;
    }
    __device__ int count() const { return n; }
    __device__ bool same(int i, f3 p, float eps) const {
        if (i >= n) return false;
        const float* s = slot + 8 * i;
        return fabsf(p.x - s[3]) < eps && fabsf(p.y - s[4]) < eps && fabsf(p.z - s[5]) < eps;
    }
};

// contact.py:354-412 (principal inertia axes of the link of the smaller geom)
template <class S>
__device__ inline void contact_axes(const DModel& m, const S& st, int b, int ga, int gb, f3 n, f3& a0, f3& a1) {
    int g = gb;
    if (m.geom_type[ga] != GT_PLANE) {
        cfptr A = &m.geom_init_aabb[24 * ga + 21];
        cfptr Bx = &m.geom_init_aabb[24 * gb + 21];
        g = (A[0] * A[1] * A[2]) < (Bx[0] * Bx[1] * Bx[2]) ? ga : gb;
    }
    m33 R = quat_R(LD4(i_quat, m.geom_link[g]), m.eps);
    int idx = 0;
    float amax = 0.f;
    for (int i = 0; i < 3; i++) {
        float a = fabsf(dot3(mk3(R.a[0][i], R.a[1][i], R.a[2][i]), n));
        if (a > amax) {
            amax = a;
            idx = i;
        }
    }
    idx = (idx + 1) % 3;
    a0 = mk3(R.a[0][idx], R.a[1][idx], R.a[2][idx]);
    a0 = normalize3(a0 - dot3(n, a0) * n);
    a1 = cross3(n, a0);
}

__device__ inline void rotate_frame(f3 pos, f4q quat, f3 cpos, f4q qr, f3& np, f4q& nq) {
    nq = qtrans(quat, qr);
    f3 rel = cpos - pos;
    np = pos - (qrot(rel, qr) - rel);
}

__device__ inline f3 qrot_fast(f3 v, f4q q) {
    f3 u = mk3(q.x, q.y, q.z);
    f3 t = 2.f * cross3(u, v);
    return v + q.w * t + cross3(u, t);
}

// utils.py:9-68
__device__ inline void closest_segments(f3 a1, f3 a2, f3 b1, f3 b2, float eps, f3& pa, f3& pb) {
    f3 da = a2 - a1, db = b2 - b1, r = a1 - b1;
    float aa = dot3(da, da), ab = dot3(da, db), bb = dot3(db, db), d = dot3(da, r), e = dot3(db, r);
    float den = aa * bb - ab * ab;
    float s = 0.f, t = 0.f;
    if (den < eps) {
        t = bb > eps ? clampf(e / bb, 0.f, 1.f) : 0.f;
    } else {
        s = clampf((ab * e - bb * d) / den, 0.f, 1.f);
        t = clampf(bb > eps ? (ab * s + e) / bb : 0.f, 0.f, 1.f);
        float sn = clampf(aa > eps ? (ab * t - d) / aa : 0.f, 0.f, 1.f);
        if (aa > eps) s = sn;
    }
    pa = a1 + s * da;
    pb = b1 + t * db;
}

// capsule_contact.py:10-99
__device__ inline bool capsule_capsule(const DModel& m, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb, f3& n, f3& cp,
                                       float& pen) {
    float ra = m.geom_data[7 * ga], ha = 0.5f * m.geom_data[7 * ga + 1];
    float rb = m.geom_data[7 * gb], hb = 0.5f * m.geom_data[7 * gb + 1];
    f3 z = mk3(0.f, 0.f, 1.f);
    f3 axa = qrot_fast(z, qa), axb = qrot_fast(z, qb);
    f3 Pa, Pb;
    closest_segments(pa - ha * axa, pa + ha * axa, pb - hb * axb, pb + hb * axb, m.eps, Pa, Pb);
    f3 diff = Pa - Pb;
    float d2 = dot3(diff, diff), cr = ra + rb;
    n = mk3(1.f, 0.f, 0.f);
    cp = zero3();
    pen = 0.f;
    if (!(d2 < cr * cr)) return false;
    float dist = sqrtf(d2);
    if (dist > m.eps) {
        n = diff / dist;
    } else {
        f3 nd = cross3(axa, axb);
        float l = dot3(nd, nd);
        if (l > m.eps) n = nd / l;
        else n = fabsf(axa.x) < 0.9f ? mk3(0.f, -axa.z, axa.y) : mk3(axa.z, 0.f, -axa.x);
    }
    pen = cr - dist;
    cp = Pa - ra * n + 0.5f * pen * n;
    return true;
}

// capsule_contact.py:102-196
__device__ inline bool sphere_capsule(const DModel& m, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb, f3& n, f3& cp,
                                      float& pen) {
    float ndir = 1.f;
    f3 sc = pa, cc = pb;
    f4q cq = qb;
    if (m.geom_type[gb] == GT_SPHERE) {
        int t = ga;
        ga = gb;
        gb = t;
        sc = pb;
        cc = pa;
        cq = qa;
        ndir = -1.f;
    }
    float sr = m.geom_data[7 * ga], cr = m.geom_data[7 * gb], ch = 0.5f * m.geom_data[7 * gb + 1];
    f3 ax = qrot_fast(mk3(0.f, 0.f, 1.f), cq);
    f3 P1 = cc - ch * ax, P2 = cc + ch * ax, seg = P2 - P1;
    float sl = dot3(seg, seg), t = 0.5f;
    if (sl > m.eps) t = clampf(dot3(sc - P1, seg) / sl, 0.f, 1.f);
    f3 diff = sc - (P1 + t * seg);
    float d2 = dot3(diff, diff), comb = sr + cr;
    n = mk3(1.f, 0.f, 0.f);
    cp = zero3();
    pen = 0.f;
    bool col = false;
    if (d2 < comb * comb) {
        col = true;
        float dist = sqrtf(d2);
        if (dist > m.eps) n = diff / dist;
        else n = fabsf(ax.x) < 0.9f ? mk3(0.f, -ax.z, ax.y) : mk3(ax.z, 0.f, -ax.x);
        pen = comb - dist;
        cp = sc - (sr - 0.5f * pen) * n;
    }
    n = n * ndir;
    return col;
}

// narrowphase.py:520-970 func_convex_convex_contact (MPR code path), in parts so that the fused kernel can run the
// perturbed detections of one pair on different lanes: convex_setup (pair constants), convex_pose (the poses of
// detection det: unperturbed for det = 0, rotated about the first contact point for det = 1..4), convex_detect (one
// detection), convex_first (after detection 0: the first contact, the warm-start cache, the perturbation axes),
// convex_correct (after detection det >= 1: contact point and normal brought back to the unperturbed frame) and
// convex_merge (the tolerance test against the pair's earlier contacts, then the append). convex_convex chains them
// in the reference's order. Each kernel calls convex_detect from one place only, so the MPR code is inlined once.
// Geometry is read from (st, b); the warm-start normal cache lives in (gs, ge); contacts go to the sink.

// what the detections of a pair need besides the poses
struct ConvexFirst {
    f3 n0, p0, ax0, ax1;
    float tolerance;
    int pair;
    bool multi;
    bool perturb;  // multi-contact pair whose detection 0 collided: detections 1..4 follow
};

__device__ inline bool convex_setup(const DModel& m, int ga, int gb, ConvexFirst& f) {
    const int ta = m.geom_type[ga], tb = m.geom_type[gb];
    f.perturb = false;
    f.multi = m.enable_multi_contact && ta != GT_SPHERE && ta != GT_ELLIPSOID && tb != GT_SPHERE && tb != GT_ELLIPSOID;
    f.tolerance = geom_tolerance(m, ga, gb, m.mc_tolerance);
    f.pair = m.collision_pair_idx[min(ga, gb) * m.G + max(ga, gb)];
    f.n0 = f.p0 = f.ax0 = f.ax1 = zero3();
    return !(ta == GT_PLANE && tb == GT_BOX);
}

template <class S>
__device__ inline void convex_pose(const DModel& m, const S& st, int b, int ga, int gb, const ConvexFirst& f,
                                   int det, f3& pa, f4q& qa, f3& pb, f4q& qb, f4q& qr) {
    pa = LD3(geoms_pos, ga);
    qa = LD4(geoms_quat, ga);
    pb = LD3(geoms_pos, gb);
    qb = LD4(geoms_quat, gb);
    qr = f4q{0.f, 0.f, 0.f, 0.f};
    if (det > 0) {
        const f3 axis = (float)(2 * (det % 2) - 1) * f.ax0 + (float)(1 - 2 * ((det / 2) % 2)) * f.ax1;
        qr = rotvec_quat(m.mc_perturbation * axis, m.eps);
        rotate_frame(pa, qa, f.p0, qr, pa, qa);
        rotate_frame(pb, qb, f.p0, qinv(qr), pb, qb);
    }
}

// one detection at poses (pa, qa), (pb, qb); detection 0 retries MPR without the warm-start normal. want_gjk is set
// when the MPR answer is to be replaced by a GJK/EPA detection at the same poses (narrowphase.py:726-734); the caller
// runs it (one lane alone, or the whole wavefront) and applies it with apply_gjk.
__device__ inline bool convex_detect(const DModel& m, const DState& gs, int ge, int ga, int gb, int pair, bool first,
                                     float tolerance, f3 pa, f4q qa, f3 pb, f4q qb, f3& n, f3& cp, float& pen,
                                     bool& want_gjk) {
    const int ta = m.geom_type[ga], tb = m.geom_type[gb];
    bool col = false;
    want_gjk = false;
    if (ta == GT_CAPSULE && tb == GT_CAPSULE) {
        col = capsule_capsule(m, ga, gb, pa, qa, pb, qb, n, cp, pen);
    } else if ((ta == GT_SPHERE && tb == GT_CAPSULE) || (ta == GT_CAPSULE && tb == GT_SPHERE)) {
        col = sphere_capsule(m, ga, gb, pa, qa, pb, qb, n, cp, pen);
    } else if (ta == GT_PLANE) {
        f3 pd = qrot(mk3(m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]), qa);
        n = -normalize3(pd);
        f3 v1 = support_driver(m, n, gb, pb, qb);
        pen = dot3(n, v1 - pa);
        cp = v1 - 0.5f * pen * n;
        col = pen > 0.f;
    } else {
        f3 nws = ld3(gs.cache_normal, pair, gs.B, ge);
        bool avail = fabsf(nws.x) > m.eps || fabsf(nws.y) > m.eps || fabsf(nws.z) > m.eps;
#pragma clang loop unroll(disable)
        for (int i_mpr = 0; i_mpr < 2; i_mpr++) {
            if (i_mpr == 1) {
                if (!(first && !col && avail)) break;
                nws = zero3();
                avail = false;
            }
            mpr_contact(m, ga, gb, nws, pa, qa, pb, qb, col, n, pen, cp);
        }
        want_gjk = prefer_gjk(m, pen, tolerance, avail);
    }
    return col;
}

// narrowphase.py:799-851: the GJK/EPA detection replaces MPR's penetration, and its normal and position when it
// found a contact
__device__ __forceinline__ bool apply_gjk(bool gcol, f3 gn, f3 gp, float gpen, f3& n, f3& cp, float& pen) {
    pen = gpen;
    if (gcol) {
        n = gn;
        cp = gp;
    }
    return gcol;
}

template <class Sink, class S>
__device__ inline void convex_first(const DModel& m, const S& st, int b, const DState& gs, int ge, int ga, int gb,
                                    bool col, f3 n, f3 cp, float pen, Sink& sink, ConvexFirst& f) {
    f.n0 = n;
    f.p0 = cp;
    if (col) {
        sink.add(ga, gb, n, cp, pen, f.pair);
        if (f.multi) {
            contact_axes(m, st, b, ga, gb, n, f.ax0, f.ax1);
            f.perturb = true;
        }
        st3(gs.cache_normal, f.pair, gs.B, ge, n);
    } else {
        st3(gs.cache_normal, f.pair, gs.B, ge, zero3());
    }
}

__device__ inline void convex_correct(const DModel& m, const ConvexFirst& f, f4q qr, f3& n, f3& cp, float& pen) {
    const float mcp = m.mc_perturbation;
    const f3 cpa = qrot((cp - 0.5f * pen * n) - f.p0, qinv(qr)) + f.p0;
    const f3 cpb = qrot((cp + 0.5f * pen * n) - f.p0, qr) + f.p0;
    cp = 0.5f * (cpa + cpb);
    f3 tw = cross3(n, f.n0);
    tw = mk3(clampf(tw.x, -mcp, mcp), clampf(tw.y, -mcp, mcp), clampf(tw.z, -mcp, mcp));
    n = n + cross3(tw, n);
    pen = dot3(n, cpb - cpa);
}

template <class Sink>
__device__ inline void convex_merge(Sink& sink, int& n_con, int ga, int gb, const ConvexFirst& f, f3 n, f3 cp,
                                    float pen) {
    bool rep = false;
    for (int i = 0; i < n_con; i++)
        if (!rep && sink.near(i, cp, f.tolerance)) rep = true;
    if (!rep && pen > -f.tolerance) {
        pen = fmaxf(pen, 0.f);
        sink.add(ga, gb, n, cp, pen, f.pair);
        n_con++;
    }
}

template <class Sink, class S>
__device__ inline void convex_convex(const DModel& m, const S& st, int b, const DState& gs, int ge, int ga, int gb,
                                     Sink& sink, float* gjk_ws) {
    ConvexFirst f;
    if (!convex_setup(m, ga, gb, f)) return;
    int n_con = 1;
#pragma clang loop unroll(disable)
    for (int det = 0; det < 5; det++) {
        if (det > 0 && !f.perturb) break;
        f3 pa, pb;
        f4q qa, qb, qr;
        convex_pose(m, st, b, ga, gb, f, det, pa, qa, pb, qb, qr);
        f3 n = zero3(), cp = zero3();
        float pen = 0.f;
        bool want_gjk;
        bool col = convex_detect(m, gs, ge, ga, gb, f.pair, det == 0, f.tolerance, pa, qa, pb, qb, n, cp, pen,
                                 want_gjk);
        if (want_gjk) {
            const GjkPair gp = make_gjk_pair(m, ga, gb, pa, qa, pb, qb);
            f3 gn = zero3(), gpos = zero3();
            float gpen = 0.f;
            const bool gcol = gjk_contact_mem(m, gp, gjk_ws, gn, gpos, gpen);
            col = apply_gjk(gcol, gn, gpos, gpen, n, cp, pen);
        }
        if (det == 0) {
            convex_first(m, st, b, gs, ge, ga, gb, col, n, cp, pen, sink, f);
        } else if (col) {
            convex_correct(m, f, qr, n, cp, pen);
            convex_merge(sink, n_con, ga, gb, f, n, cp, pen);
        }
    }
}

// box_contact.py:25-97 func_plane_box_contact
template <class Sink, class S>
__device__ inline void plane_box(const DModel& m, const S& st, int b, int ga, int gb, Sink& sink) {
    f3 pa = LD3(geoms_pos, ga), pb = LD3(geoms_pos, gb);
    f4q qa = LD4(geoms_quat, ga), qb = LD4(geoms_quat, gb);
    int pair = m.collision_pair_idx[min(ga, gb) * m.G + max(ga, gb)];
    f3 n = -normalize3(qrot(mk3(m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]), qa));
    f3 v1 = support_driver(m, n, gb, pb, qb);
    float pen = dot3(n, v1 - pa);
    if (!(pen > 0.f)) return;
    f3 cp = v1 - 0.5f * pen * n;
    sink.add(ga, gb, n, cp, pen, pair);
    if (!m.enable_multi_contact) return;
    int n_con = 1;
    float tol = geom_tolerance(m, ga, gb, m.mc_tolerance);
    cfptr d = &m.geom_data[7 * gb];
    for (int c = 0; c < 8; c++) {
        if (n_con >= m.NCPP) break;
        f3 loc = mk3(((c & 4) ? 0.5f : -0.5f) * d[0], ((c & 2) ? 0.5f : -0.5f) * d[1], ((c & 1) ? 0.5f : -0.5f) * d[2]);
        f3 corner = tr_point(loc, pb, qb);
        float p = dot3(n, corner - pa);
        if (p > 0.f) {
            f3 c2 = corner - 0.5f * p * n;
            if (norm3(c2 - cp) > tol) {
                sink.add(ga, gb, n, c2, p, pair);
                n_con++;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------------- box-box
// box_contact.py:100-785 func_box_box_contact: separating-axis test over the 15 axes of two boxes, then the deepest
// feature clipped against the reference face (face contact) or the two edges clipped against each other's faces
// (edge contact). Candidate points go straight to the sink in the order the reference collects them; the reference's
// dedup (every coordinate within EPS of a contact this pair already added) and its n_contacts_per_pair limit apply as
// they are emitted. Components are picked with selects (no dynamically indexed private arrays).
__device__ __forceinline__ float mel(const m33& A, int i, int j) {
    const f3 r = i == 0 ? mk3(A.a[0][0], A.a[0][1], A.a[0][2])
                        : (i == 1 ? mk3(A.a[1][0], A.a[1][1], A.a[1][2]) : mk3(A.a[2][0], A.a[2][1], A.a[2][2]));
    return comp(r, j);
}
__device__ __forceinline__ f3 mrow3(const m33& A, int i) {
    return i == 0 ? mk3(A.a[0][0], A.a[0][1], A.a[0][2])
                  : (i == 1 ? mk3(A.a[1][0], A.a[1][1], A.a[1][2]) : mk3(A.a[2][0], A.a[2][1], A.a[2][2]));
}
// contact.py:16-31 rotaxis / rotmatx
__device__ __forceinline__ f3 rotaxis3(f3 v, int i0, int i1, int i2, float f0, float f1, float f2) {
    return mk3(comp(v, i0) * f0, comp(v, i1) * f1, comp(v, i2) * f2);
}
__device__ __forceinline__ m33 rotmatx3(const m33& A, int i0, int i1, int i2, float f0, float f1, float f2) {
    const f3 r0 = mrow3(A, i0) * f0, r1 = mrow3(A, i1) * f1, r2 = mrow3(A, i2) * f2;
    return m33{{{r0.x, r0.y, r0.z}, {r1.x, r1.y, r1.z}, {r2.x, r2.y, r2.z}}};
}
// box_contact.py:241-303 / :488-520: the frame that puts face q (0..5) of the reference box on +z
__device__ __forceinline__ void box_face_frame(int q, m33& R, int& i0, int& i1, int& i2, float& f0, float& f1, float& f2) {
    R = m33{{{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}}};
    i0 = 0, i1 = 1, i2 = 2;
    f0 = f1 = f2 = 1.f;
    if (q == 0) R.a[0][2] = -1.f, R.a[1][1] = 1.f, R.a[2][0] = 1.f, i0 = 2, f0 = -1.f, i2 = 0;
    else if (q == 1) R.a[0][0] = 1.f, R.a[1][2] = -1.f, R.a[2][1] = 1.f, i1 = 2, f1 = -1.f, i2 = 1;
    else if (q == 2) R.a[0][0] = 1.f, R.a[1][1] = 1.f, R.a[2][2] = 1.f;
    else if (q == 3) R.a[0][2] = 1.f, R.a[1][1] = 1.f, R.a[2][0] = -1.f, i0 = 2, i2 = 0, f2 = -1.f;
    else if (q == 4) R.a[0][0] = 1.f, R.a[1][2] = 1.f, R.a[2][1] = -1.f, i1 = 2, i2 = 1, f2 = -1.f;
    else R.a[0][0] = -1.f, R.a[1][1] = 1.f, R.a[2][2] = -1.f, f0 = -1.f, f2 = -1.f;
}

// adds one candidate (local point with its height already in place, distance dist) as the reference's add loop does
template <class Sink>
struct BoxEmit {
    Sink& sink;
    int ga, gb, pair, start, added, cap;
    f3 base, nrm;
    m33 R;
    float hz, eps;
    __device__ __forceinline__ void operator()(f3 pt, float dist) {
        if (added >= cap) return;
        pt.z = pt.z + hz;
        const f3 cpos = base + mv33(R, pt);
        bool valid = true;
        for (int j = 0; j < added; j++)
            if (sink.same(start + j, cpos, eps)) valid = false;
        if (valid) {
            sink.add(ga, gb, nrm, cpos, -dist, pair);
            added++;
        }
    }
};

template <class Sink, class S>
__device__ inline void box_box(const DModel& m, const S& st, int b, int ga, int gb, Sink& sink) {
    const float EPS = m.eps, margin = 0.f, margin2 = margin * margin;
    const int pair = m.collision_pair_idx[min(ga, gb) * m.G + max(ga, gb)];
    const f3 pos1 = LD3(geoms_pos, ga), pos2 = LD3(geoms_pos, gb);
    const f3 size1 = mk3(m.geom_data[7 * ga] / 2.f, m.geom_data[7 * ga + 1] / 2.f, m.geom_data[7 * ga + 2] / 2.f);
    const f3 size2 = mk3(m.geom_data[7 * gb] / 2.f, m.geom_data[7 * gb + 1] / 2.f, m.geom_data[7 * gb + 2] / 2.f);
    const m33 mat1 = quat_R(LD4(geoms_quat, ga), EPS), mat2 = quat_R(LD4(geoms_quat, gb), EPS);
    const f3 pos21 = mv33(tr33(mat1), pos2 - pos1);
    const f3 pos12 = mv33(tr33(mat2), pos1 - pos2);
    const m33 rot = mm33(tr33(mat1), mat2), rott = tr33(rot);
    m33 rotabs, rottabs;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) rotabs.a[i][j] = fabsf(rot.a[i][j]), rottabs.a[i][j] = fabsf(rott.a[i][j]);
    const f3 plen2 = mv33(rotabs, size2);
    const f3 plen1 = mv33(tr33(rotabs), size1);
    float penetration = margin;
#pragma unroll
    for (int i = 0; i < 3; i++) penetration = penetration + size1[i] * 3.f + size2[i] * 3.f;
    int code = -1, cle1 = 0, cle2 = 0;
    bool is_return = false, in_ = false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float c1 = -fabsf(pos21[i]) + size1[i] + plen2[i];
        const float c2 = -fabsf(pos12[i]) + size2[i] + plen1[i];
        if (c1 < -margin || c2 < -margin) is_return = true;
        if (c1 < penetration) penetration = c1, code = i + 3 * (pos21[i] < 0.f ? 1 : 0);
        if (c2 < penetration) penetration = c2, code = i + 3 * (pos12[i] < 0.f ? 1 : 0) + 6;
    }
    f3 clnorm = zero3();
#pragma unroll
    for (int i = 0; i < 3; i++) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float rj0 = rott.a[j][0], rj1 = rott.a[j][1], rj2 = rott.a[j][2];
            f3 t = i == 0 ? mk3(0.f, -rj2, rj1) : (i == 1 ? mk3(rj2, 0.f, -rj0) : mk3(-rj1, rj0, 0.f));
            const float c1 = norm3(t);
            t = mk3(t.x / c1, t.y / c1, t.z / c1);
            if (c1 >= EPS) {
                const float c2 = dot3(pos21, t);
                float c3 = 0.f;
#pragma unroll
                for (int k = 0; k < 3; k++)
                    if (k != i) c3 = c3 + size1[k] * fabsf(t[k]);
#pragma unroll
                for (int k = 0; k < 3; k++)
                    if (k != j) c3 = c3 + size2[k] * rotabs.a[i][3 - k - j] / c1;
                c3 = c3 - fabsf(c2);
                if (c3 < -margin) is_return = true;
                if (c3 < penetration * (1.f - 1e-12f)) {
                    penetration = c3;
                    cle1 = 0;
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        if (k != i && ((t[k] > 0.f) != (c2 < 0.f))) cle1 += 1 << k;
                    cle2 = 0;
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        if (k != j && (((rot.a[i][3 - k - j] > 0.f) != (c2 < 0.f)) != ((k - j + 3) % 3 == 1)))
                            cle2 += 1 << k;
                    code = 12 + i * 3 + j;
                    clnorm = t;
                    in_ = c2 < 0.f;
                }
            }
        }
    }
    if (code == -1 || is_return) return;
    m33 rotmore;
    int i0, i1, i2;
    float f0, f1, f2;
    BoxEmit<Sink> emit{sink, ga, gb, pair, sink.count(), 0, m.NCPP, zero3(), zero3(), rotmore, 0.f, EPS};
    if (code < 12) {
        // face contact (box_contact.py:241-457)
        const int q1 = code % 6, q2 = code / 6;
        box_face_frame(q1, rotmore, i0, i1, i2, f0, f1, f2);
        m33 r;
        f3 p, t1, sz;
        if (q2) {
            r = mm33(rotmore, tr33(rot));
            p = rotaxis3(pos12, i0, i1, i2, f0, f1, f2);
            t1 = rotaxis3(size2, i0, i1, i2, f0, f1, f2);
            sz = size1;
        } else {
            r = rotmatx3(rot, i0, i1, i2, f0, f1, f2);
            p = rotaxis3(pos21, i0, i1, i2, f0, f1, f2);
            t1 = rotaxis3(size1, i0, i1, i2, f0, f1, f2);
            sz = size2;
        }
        const m33 rt = tr33(r);
        const f3 ss = mk3(fabsf(t1.x), fabsf(t1.y), fabsf(t1.z));
        const float lx = ss.x, ly = ss.y, hz = ss.z;
        p.z = p.z - hz;
        int clcorner = 0;
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (r.a[2][i] < 0.f) clcorner += 1 << i;
        f3 lp = p;
#pragma unroll
        for (int i = 0; i < 3; i++)
            lp = lp + mk3(rt.a[i][0], rt.a[i][1], rt.a[i][2]) * sz[i] * ((clcorner & (1 << i)) ? 1.f : -1.f);
        const f3 pts0 = lp;
        f3 pts1 = zero3(), pts2 = zero3();
        int mm = 1;
#pragma unroll
        for (int i = 0; i < 3; i++)
            if (fabsf(r.a[2][i]) < 0.5f) {
                const f3 e = mk3(rt.a[i][0], rt.a[i][1], rt.a[i][2]) * sz[i] * ((clcorner & (1 << i)) ? -2.f : 2.f);
                if (mm == 1) pts1 = e;
                else pts2 = e;
                mm++;
            }
        const f3 pts3 = pts0 + pts1, pts4 = pts0 + pts2, pts5 = pts3 + pts2;
        const m33 rr = mm33(q2 ? mat2 : mat1, tr33(rotmore));
        const float sg = q2 ? -1.f : 1.f;
        emit.base = q2 ? pos2 : pos1;
        emit.nrm = -mk3(sg * rr.a[0][2], sg * rr.a[1][2], sg * rr.a[2][2]);
        emit.R = rr;
        emit.hz = hz;
        // a candidate at or below the face: its distance is its height, halved (box_contact.py:414-434)
        auto cand = [&](f3 pt) {
            if (pt.z <= margin) {
                pt.z = pt.z * 0.5f;
                emit(pt, pt.z);
            }
        };
        auto clip = [&](f3 A, f3 D) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                const float a = A[q], bb = D[q], c = A[1 - q], d = D[1 - q];
                if (fabsf(bb) > EPS) {
#pragma unroll
                    for (int jj = 0; jj < 2; jj++) {
                        const float l = ss[q] * (float)(2 * jj - 1);
                        const float c1 = (l - a) / bb;
                        if (0.f <= c1 && c1 <= 1.f) {
                            const float c2 = c + d * c1;
                            if (fabsf(c2) <= ss[1 - q]) cand(A + D * c1);
                        }
                    }
                }
            }
        };
        if (mm > 1) clip(pts0, pts1);
        if (mm > 2) {
            clip(pts0, pts2);
            clip(pts3, pts2);
            clip(pts4, pts1);
        }
        const float a = pts1.x, bb = pts2.x, c = pts1.y, d = pts2.y;
        const float c1 = a * d - bb * c;
        if (mm > 2) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float llx = (i / 2) ? lx : -lx, lly = (i % 2) ? ly : -ly;
                const float x = llx - pts0.x, y = lly - pts0.y;
                const float u = (x * d - y * bb) / c1, v = (y * a - x * c) / c1;
                if (0.f < u && u < 1.f && 0.f < v && v < 1.f) cand(mk3(llx, lly, pts0.z + u * pts1.z + v * pts2.z));
            }
        }
        const int nv = 1 << (mm - 1);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (i < nv) {
                const f3 t = i == 0 ? pts0 : (i == 1 ? pts3 : (i == 2 ? pts4 : pts5));
                if (!(i && (t.x <= -lx || t.x >= lx || t.y <= -ly || t.y >= ly))) cand(t);
            }
        }
        return;
    }
    // edge-edge contact (box_contact.py:458-785)
    const int ce = code - 12;
    const int q1 = ce / 3, q2 = ce % 3;
    int ax1 = q2 == 1 ? 0 : 1, ax2 = q2 == 0 ? 2 : (q2 == 1 ? 2 : 0);
    int pax1 = q1 == 1 ? 0 : 1, pax2 = q1 == 0 ? 2 : (q1 == 1 ? 2 : 0);
    if (mel(rotabs, q1, ax1) < mel(rotabs, q1, ax2)) {
        ax1 = ax2;
        ax2 = 3 - q2 - ax1;
    }
    if (mel(rottabs, q2, pax1) < mel(rottabs, q2, pax2)) {
        pax1 = pax2;
        pax2 = 3 - q1 - pax1;
    }
    const int clface = (cle1 & (1 << pax2)) ? pax2 : pax2 + 3;
    box_face_frame(clface, rotmore, i0, i1, i2, f0, f1, f2);
    f3 p = rotaxis3(pos21, i0, i1, i2, f0, f1, f2);
    const f3 rnorm = rotaxis3(clnorm, i0, i1, i2, f0, f1, f2);
    const m33 r = rotmatx3(rot, i0, i1, i2, f0, f1, f2);
    const f3 t1s = mv33(tr33(rotmore), size1);
    const f3 sz = mk3(fabsf(t1s.x), fabsf(t1s.y), fabsf(t1s.z));
    const m33 rt = tr33(r);
    const float lx = sz.x, ly = sz.y, hz = sz.z;
    p.z = p.z - hz;
    const float e1 = (cle2 & (1 << ax1)) ? 1.f : -1.f, e2 = (cle2 & (1 << ax2)) ? 1.f : -1.f;
    const f3 r_ax1 = mrow3(rt, ax1), r_ax2 = mrow3(rt, ax2), r_q2 = mrow3(rt, q2);
    const float s_ax1 = comp(size2, ax1), s_ax2 = comp(size2, ax2), s_q2 = comp(size2, q2);
    f3 u0 = p;
    u0 = u0 + r_ax1 * s_ax1 * e1;
    u0 = u0 + r_ax2 * s_ax2 * e2;
    f3 u1 = u0;
    u0 = u0 + r_q2 * s_q2;
    u1 = u1 - r_q2 * s_q2;
    f3 u2 = p;
    u2 = u2 + r_ax1 * s_ax1 * (-e1);
    u2 = u2 + r_ax2 * s_ax2 * e2;
    f3 u3 = u2;
    u2 = u2 + r_q2 * s_q2;
    u3 = u3 - r_q2 * s_q2;
    const f3 axi0 = u0, axi1 = u1 - u0, axi2 = u2 - u0;
    if (fabsf(rnorm.z) < EPS) return;
    const float innorm = (1.f / rnorm.z) * (in_ ? -1.f : 1.f);
    // pu: the points as they were; q: projected along the normal onto the face plane
    const f3 pu0 = u0, pu1 = u1, pu2 = u2, pu3 = u3;
    const f3 q0 = u0 + (-u0.z / rnorm.z) * rnorm, q1p = u1 + (-u1.z / rnorm.z) * rnorm;
    const f3 q2p = u2 + (-u2.z / rnorm.z) * rnorm, q3p = u3 + (-u3.z / rnorm.z) * rnorm;
    const f3 pts0 = q0, pts1 = q1p - q0, pts2 = q2p - q0;
    const m33 rr = mm33(mat1, tr33(rotmore));
    const f3 nr = mv33(rr, rnorm);
    emit.base = pos1;
    emit.nrm = -((in_ ? -1.f : 1.f) * nr);
    emit.R = rr;
    emit.hz = hz;
    int n = 0;
    auto eclip = [&](f3 A, f3 D, f3 UA, f3 UD) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const float a = A[q], bb = D[q], c = A[1 - q], d = D[1 - q];
            if (fabsf(bb) > EPS) {
#pragma unroll
                for (int jj = 0; jj < 2; jj++) {
                    const float l = sz[q] * (float)(2 * jj - 1);
                    const float c1 = (l - a) / bb;
                    if (0.f <= c1 && c1 <= 1.f) {
                        const float c2 = c + d * c1;
                        if (fabsf(c2) <= sz[1 - q] && (UA.z + UD.z * c1) * innorm <= margin) {
                            f3 pt = UA * 0.5f + c1 * 0.5f * UD;
                            if (q == 0) pt.x = pt.x + 0.5f * l, pt.y = pt.y + 0.5f * c2;
                            else pt.y = pt.y + 0.5f * l, pt.x = pt.x + 0.5f * c2;
                            n++;
                            emit(pt, pt.z * innorm * 2.f);
                        }
                    }
                }
            }
        }
    };
    eclip(pts0, pts1, axi0, axi1);
    eclip(pts0, pts2, axi0, axi2);
    eclip(pts0 + pts1, pts2, axi0 + axi1, axi2);
    eclip(pts0 + pts2, pts1, axi0 + axi2, axi1);
    const int nl = n;
    {
        const float a = pts1.x, bb = pts2.x, c = pts1.y, d = pts2.y;
        const float c1 = a * d - bb * c;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float llx = (i / 2) ? lx : -lx, lly = (i % 2) ? ly : -ly;
            const float x = llx - pts0.x, y = lly - pts0.y;
            float u = (x * d - y * bb) / c1, v = (y * a - x * c) / c1;
            bool skip;
            if (nl == 0) skip = (u < 0.f || u > 1.f) && (v < 0.f || v > 1.f);
            else skip = u < 0.f || u > 1.f || v < 0.f || v > 1.f;
            if (skip) continue;
            u = fminf(fmaxf(u, 0.f), 1.f);
            v = fminf(fmaxf(v, 0.f), 1.f);
            const f3 t1 = pu0 * (1.f - u - v) + pu1 * u + pu2 * v;
            f3 pt = mk3(llx, lly, 0.f);
            const f3 t2 = pt - t1;
            const float cc = dot3(t2, t2);
            if (!(t1.z > 0.f && cc > margin2)) {
                pt = pt + t1;
                pt = pt * 0.5f;
                n++;
                emit(pt, sqrtf(cc) * (t1.z < 0.f ? -1.f : 1.f));
            }
        }
    }
    const int nf = n;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const f3 qi = i == 0 ? q0 : (i == 1 ? q1p : (i == 2 ? q2p : q3p));
        const f3 pui = i == 0 ? pu0 : (i == 1 ? pu1 : (i == 2 ? pu2 : pu3));
        const float x = qi.x, y = qi.y;
        bool skip;
        if (nl == 0) skip = nf != 0 && (x < -lx || x > lx) && (y < -ly || y > ly);
        else skip = x < -lx || x > lx || y < -ly || y > ly;
        if (skip) continue;
        float cc = 0.f;
        if (x < -sz.x) cc = cc + (x + sz.x) * (x + sz.x);
        else if (x > sz.x) cc = cc + (x - sz.x) * (x - sz.x);
        if (y < -sz.y) cc = cc + (y + sz.y) * (y + sz.y);
        else if (y > sz.y) cc = cc + (y - sz.y) * (y - sz.y);
        cc = cc + (pui.z * innorm) * (pui.z * innorm);
        if (pui.z > 0.f && cc > margin2) continue;
        f3 t1 = mk3(x * 0.5f, y * 0.5f, 0.f);
        if (x < -sz.x) t1.x = -sz.x * 0.5f;
        else if (x > sz.x) t1.x = sz.x * 0.5f;
        if (y < -sz.y) t1.y = -sz.y * 0.5f;
        else if (y > sz.y) t1.y = sz.y * 0.5f;
        t1 = t1 + pui * 0.5f;
        n++;
        emit(t1, sqrtf(cc) * (pui.z < 0.f ? -1.f : 1.f));
    }
}

// narrowphase.py:1988-2060 + :2135-2189
__device__ inline bool convex_pair(const DModel& m, int ta, int tb, int ga, int gb) {
    return m.geom_is_convex[ga] && m.geom_is_convex[gb] && tb != GT_TERRAIN &&
           !(m.box_box_detection && ta == GT_BOX && tb == GT_BOX) && !(ta == GT_PLANE && tb == GT_BOX);
}

template <class S>
__device__ inline void dev_narrowphase(const DModel& m, const S& st, int b) {
    PlaneSink sink{m, st, b};
    float* gjk_ws = st.wk + (size_t)b * m.work_stride + m.off_gjk;
    int nb = SF(n_broad, 0);
    for (int i = 0; i < nb; i++) {
        int ga = SF(broad_pairs, 2 * i), gb = SF(broad_pairs, 2 * i + 1);
        if (m.geom_type[ga] > m.geom_type[gb]) {
            int t = ga;
            ga = gb;
            gb = t;
        }
        int ta = m.geom_type[ga], tb = m.geom_type[gb];
        if (convex_pair(m, ta, tb, ga, gb)) convex_convex(m, st, b, st, b, ga, gb, sink, gjk_ws);
    }
    for (int i = 0; i < nb; i++) {
        int ga = SF(broad_pairs, 2 * i), gb = SF(broad_pairs, 2 * i + 1);
        if (m.geom_type[ga] > m.geom_type[gb]) {
            int t = ga;
            ga = gb;
            gb = t;
        }
        if (m.geom_type[ga] == GT_PLANE && m.geom_type[gb] == GT_BOX) plane_box(m, st, b, ga, gb, sink);
        if (m.box_box_detection && m.geom_type[ga] == GT_BOX && m.geom_type[gb] == GT_BOX) box_box(m, st, b, ga, gb, sink);
    }
}

}  // namespace rs
