// GJK/EPA contact detection: the hand-over target of the MPR path for deep or poorly warm-started overlaps
// (narrowphase.py:726-851). Restates the reference's default (non-MuJoCo-compatible) path:
// func_gjk_contact -> func_safe_gjk -> func_safe_epa_init -> func_safe_epa (collider/gjk.py:170-446, :1209-1915;
// epa.py:273-434, :969-1432; gjk_utils.py; gjk_support.py; support_field.py:121-395). The polygon-clipping
// multi-contact of multi_contact.py is reached only with enable_mujoco_multi_contact, which the reference fixes to False
// (gjk.py:51), so one GJK detection gives at most one contact.
//
// The algorithm is written once against a storage policy for the EPA polytope:
//  * MemStore  - the polytope in the environment's slice of the work field at the reference's capacities; one lane
//                runs a detection alone (the per-stage kernels, one lane per environment), or every lane of a
//                wavefront runs it with the same values;
//  * LdsStore  - the polytope in the wavefront's LDS at capacities that cover the detections met in practice; every
//                lane runs the sequential parts with the same values and reads single entries as LDS broadcasts, and
//                the scans over faces and vertices (nearest face, farthest vertex from a new face, repeated support
//                ids, the exhaustive vertex-pair search) run one entry per lane with a DPP reduction; a mesh's support
//                lookup runs its four table candidates, for both geoms, on eight lanes at once. A detection that would
//                outgrow it runs again from the start with MemStore, so every result is the reference algorithm's.
// The fused kernel hands one detection at a time to the whole wavefront.
// Capacities are the reference's (GJKState sizes, utils/array_class.py:1069-1115): 50 GJK and 50 EPA iterations,
// 5 + 50 polytope vertices, 6 * 50 faces, 6 + 50 horizon edges and a horizon stack of 3 * 300 entries. Where the
// reference would write past one of them (it does not check), the EPA stops as for a failed horizon.
#pragma once
#include "rs_device.h"
#include "rs_wave.h"

namespace rs {

constexpr int kGjkIters = 50, kEpaIters = 50;
constexpr int kEpaFaces = 6 * kEpaIters;
constexpr int kEpaVerts = 5 + kEpaIters;
constexpr int kEpaHorizon = 6 + kEpaIters;
constexpr int kEpaStack = 3 * kEpaFaces;
constexpr int kGjkSeparated = 0, kGjkIntersect = 1, kGjkNumError = 2;  // constants.py GJK_RETURN_CODE
constexpr int kRcOk = 0, kRcFail = 1;                                  // constants.py RETURN_CODE

// gjk.py:76-106 get_gjk_info (FP32 values, computed as the reference computes them)
struct GjkInfo {
    static __device__ __forceinline__ float fmin_() { return 1e-15f; }
    static __device__ __forceinline__ float fmin_sq() { return 1e-15f * 1e-15f; }
    static __device__ __forceinline__ float fmax_() { return 1e15f; }
    static __device__ __forceinline__ float fmax_sq() { return 1e15f * 1e15f; }
    static __device__ __forceinline__ float tol() { return 1e-6f; }
    static __device__ __forceinline__ float degeneracy_sq() { return 1e-5f * 1e-5f; }
    static __device__ __forceinline__ float max_rel_reproj() { return 1e-4f; }
    static __device__ __forceinline__ float max_abs_reproj() { return 1e-4f; }
};

// readlane helpers (v's value in lane l, l wave-uniform)
__device__ __forceinline__ int rl_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float rl_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
__device__ __forceinline__ f3 rl3(f3 v, int l) { return mk3(rl_f(v.x, l), rl_f(v.y, l), rl_f(v.z, l)); }
__device__ __forceinline__ f4q rl4(f4q q, int l) { return f4q{rl_f(q.w, l), rl_f(q.x, l), rl_f(q.y, l), rl_f(q.z, l)}; }

// ------------------------------------------------------------------------------------------------- support

// What the detectors read of a geom, loaded once per detection.
struct GInfo {
    int type, sup, vstart, vend;  // geom type, support-table start (-1: none), vertex range
    float d0, d1, d2;             // geom_data[0..2] (sizes)
};
__device__ __forceinline__ GInfo ginfo(const DModel& m, int g) {
    GInfo r;
    r.type = m.geom_type[g];
    r.sup = m.geom_support_start[g];
    r.vstart = m.geom_vert_start[g];
    r.vend = m.geom_vert_end[g];
    r.d0 = m.geom_data[7 * g];
    r.d1 = m.geom_data[7 * g + 1];
    r.d2 = m.geom_data[7 * g + 2];
    return r;
}

// support_field.py:139-181 _func_support_mesh with the id of the chosen vertex (support_vid); start: the geom's
// support-table start, vstart: its first vertex (a cell's point is the vertex whose id the cell holds)
__device__ __forceinline__ f3 support_mesh_id(const DModel& m, f3 d, int start, int vstart, int& vid) {
    const int res = 180;
    const float pi = 3.14159265358979323846f;
    const float theta = atan2f(d.y, d.x);
    const float phi = acosf(clampf(d.z, -1.f, 1.f));
    const float ii = (theta + pi) / pi / 2.f * (float)res;
    const float jj = phi / pi * (float)res;
    float dmax = -1e20f;
    f3 v = zero3();
    vid = 0;
    for (int i4 = 0; i4 < 4; i4++) {
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
        i = min(max(i, 0), res - 1);
        const int id = m.support_vid[(size_t)(start + i * res + j)];
        const f3 p = mvec3(m.verts, vstart + id);
        const float dt = dot3(p, d);
        if (dt > dmax) {
            v = p;
            dmax = dt;
            vid = id;
        }
    }
    return v;
}

// gjk_support.py:64-106 support_driver (no sphere shrinking): world support point along d and the id of the vertex it
// came from (-1 for smooth geoms)
__device__ __forceinline__ f3 gjk_support_geom(const DModel& m, f3 d, const GInfo& gi, f3 pos, f4q quat, int& vid) {
    const int t = gi.type;
    vid = -1;
    if (t == GT_SPHERE) return pos + d * gi.d0;
    if (t == GT_ELLIPSOID) {
        const f3 dl = qrot_inv(d, quat);
        const float a = gi.d0, b = gi.d1, c = gi.d2;
        const float nrm = sqrtf(a * a * dl.x * dl.x + b * b * dl.y * dl.y + c * c * dl.z * dl.z);
        const f3 sl = mk3(a * a * dl.x / nrm, b * b * dl.y / nrm, c * c * dl.z / nrm);
        return pos + qrot(sl, quat);
    }
    if (t == GT_CAPSULE) {
        const f3 axis = qrot(mk3(0.f, 0.f, 1.f), quat);
        const float side = dot3(d, axis) < 0.f ? -1.f : 1.f;
        const f3 ep = pos + (0.5f * gi.d1) * side * axis;
        return ep + d * gi.d0;
    }
    if (t == GT_BOX) {
        const f3 db = qrot_inv(d, quat);
        const f3 vl = mk3((db.x < 0.f ? -1.f : 1.f) * gi.d0 * 0.5f, (db.y < 0.f ? -1.f : 1.f) * gi.d1 * 0.5f,
                          (db.z < 0.f ? -1.f : 1.f) * gi.d2 * 0.5f);
        vid = (vl.x > 0.f ? 1 : 0) + (vl.y > 0.f ? 2 : 0) + (vl.z > 0.f ? 4 : 0) + gi.vstart;
        return tr_point(vl, pos, quat);
    }
    if (gi.sup < 0) return pos;
    const f3 dm = qrot(d, qinv(quat));
    return tr_point(support_mesh_id(m, dm, gi.sup, gi.vstart, vid), pos, quat);
}

// support_field.py:312-395: a box has 2^(zero components of the local direction) support points; the mesh count
// compares against a maximum that is never raised, so it is always 1
__device__ __forceinline__ int count_support_geom(f3 d, const GInfo& gi, f4q quat) {
    if (gi.type == GT_BOX) {
        const f3 db = qrot_inv(d, quat);
        const int z = (db.x == 0.f) + (db.y == 0.f) + (db.z == 0.f);
        return 1 << z;
    }
    return 1;
}

// gjk.py:1674-1708 func_get_discrete_geom_vertex
__device__ __forceinline__ f3 discrete_vertex(const DModel& m, const GInfo& gi, f3 pos, f4q quat, int iv) {
    f3 vl = zero3();
    if (gi.type == GT_BOX) {
        vl = mk3(((iv & 1) ? 1.f : -1.f) * gi.d0 * 0.5f, ((iv & 2) ? 1.f : -1.f) * gi.d1 * 0.5f,
                 ((iv & 4) ? 1.f : -1.f) * gi.d2 * 0.5f);
    } else if (gi.type == GT_MESH) {
        vl = mvec3(m.verts, gi.vstart + iv);
    }
    return tr_point(vl, pos, quat);
}

__device__ __forceinline__ bool is_discrete(const GInfo& gi) { return gi.type == GT_MESH || gi.type == GT_BOX; }

// the two geoms of one detection at the poses it uses
struct GjkPair {
    int ga, gb;
    f3 pa, pb;
    f4q qa, qb;
    GInfo ia, ib;
};
__device__ __forceinline__ GjkPair make_gjk_pair(const DModel& m, int ga, int gb, f3 pa, f4q qa, f3 pb, f4q qb) {
    GjkPair p;
    p.ga = ga;
    p.gb = gb;
    p.pa = pa;
    p.qa = qa;
    p.pb = pb;
    p.qb = qb;
    p.ia = ginfo(m, ga);
    p.ib = ginfo(m, gb);
    return p;
}

// a support point of the Minkowski difference and the points and vertex ids it came from
struct SVert {
    f3 o1, o2, mk;
    int id1, id2;
};

__device__ __forceinline__ f3 sel3(bool c, f3 a, f3 b) { return mk3(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z); }
__device__ __forceinline__ SVert sel(bool c, const SVert& a, const SVert& b) {
    SVert r;
    r.o1 = sel3(c, a.o1, b.o1);
    r.o2 = sel3(c, a.o2, b.o2);
    r.mk = sel3(c, a.mk, b.mk);
    r.id1 = c ? a.id1 : b.id1;
    r.id2 = c ? a.id2 : b.id2;
    return r;
}

// gjk_support.py:109-176 func_support
__device__ __forceinline__ SVert support_pair_ids(const DModel& m, const GjkPair& p, f3 d) {
    SVert s;
    s.o1 = gjk_support_geom(m, d, p.ia, p.pa, p.qa, s.id1);
    s.o2 = gjk_support_geom(m, -d, p.ib, p.pb, p.qb, s.id2);
    s.mk = s.o1 - s.o2;
    return s;
}

// Both support points of a pair on one wavefront: lanes 0-31 work for geom a along d, lanes 32-63 for geom b along
// -d. A mesh's four table candidates (support_field.py:139-181) are looked up by lanes 0-3 (32-35) side by side; the
// first candidate with the largest projection wins, as in the sequential loop. Every lane gets the same result.
__device__ __forceinline__ SVert support_pair_wave(const DModel& m, const GjkPair& p, f3 d, int lane) {
    const bool second = lane >= kWave / 2;
    const int base = second ? kWave / 2 : 0;
    const GInfo gi = second ? p.ib : p.ia;
    const f3 dir = second ? -d : d;
    const f3 pos = second ? p.pb : p.pa;
    const f4q quat = second ? p.qb : p.qa;
    f3 sp;
    int vid = -1, src = base;
    if (gi.type == GT_MESH && gi.sup >= 0) {
        const int res = 180;
        const float pi = 3.14159265358979323846f;
        const f3 dm = qrot(dir, qinv(quat));
        const float theta = atan2f(dm.y, dm.x);
        const float phi = acosf(clampf(dm.z, -1.f, 1.f));
        const float ii = (theta + pi) / pi / 2.f * (float)res;
        const float jj = phi / pi * (float)res;
        const int i4 = lane & 3;
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
        i = min(max(i, 0), res - 1);
        const int id = m.support_vid[(size_t)(gi.sup + i * res + j)];
        const f3 v = mvec3(m.verts, gi.vstart + id);
        const float dt = dot3(v, dm);
        const bool ok = dt > -1e20f;
        const float key = ok ? dt : -__int_as_float(0x7f800000);
        // largest key of the quad, then its first lane
        float q = fmaxf(key, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(key), __float_as_int(key),
                                                                          0xb1, 0xf, 0xf, false)));  // [1,0,3,2]
        q = fmaxf(q, __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(q), __float_as_int(q), 0x4e, 0xf, 0xf,
                                                                false)));  // [2,3,0,1]
        const uint64_t win = wave_ballot(ok && key == q && (lane & (kWave / 2 - 1)) < 4) >> base;
        const bool found = (win & 0xf) != 0;
        src = found ? base + __builtin_ctzll(win & 0xf) : base;
        sp = tr_point(found ? v : zero3(), pos, quat);
        vid = found ? id : 0;
    } else {
        sp = gjk_support_geom(m, dir, gi, pos, quat, vid);
    }
    SVert r;
    const int la = __builtin_amdgcn_readfirstlane(src);
    const int lb = __builtin_amdgcn_readlane(src, kWave / 2);
    r.o1 = rl3(sp, la);
    r.id1 = rl_i(vid, la);
    r.o2 = rl3(sp, lb);
    r.id2 = rl_i(vid, lb);
    r.mk = r.o1 - r.o2;
    return r;
}

// --------------------------------------------------------------------------------------------------- simplex
// The GJK simplex (at most 4 vertices) is held in registers; on a wavefront every lane holds the same copy. Every
// index into it is resolved against compile-time slots, so the arrays stay in registers.
struct Simplex {
    SVert v[4];
    int n;
    __device__ __forceinline__ SVert get(int i) const {
        SVert r = v[0];
#pragma unroll
        for (int k = 1; k < 4; k++)
            if (i == k) r = v[k];
        return r;
    }
    // every slot is rewritten through a select, so that no store depends on a branch (a conditional store to a slot
    // picked at run time would be merged into one store through a computed address)
    __device__ __forceinline__ void set(int i, const SVert& s) {
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = sel(i == k, s, v[k]);
    }
};

// gjk.py:1449-1468
__device__ __forceinline__ bool simplex_duplicate(const Simplex& s, int id1, int id2) {
    bool dup = false;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (i < s.n && id1 != -1 && s.v[i].id1 == id1 && id2 != -1 && s.v[i].id2 == id2) dup = true;
    }
    return dup;
}

// gjk.py:1513-1548
__device__ __forceinline__ bool is_colinear(f3 v1, f3 v2, f3 v3) {
    const f3 e1 = v2 - v1, e2 = v3 - v1;
    const f3 n = cross3(e1, e2);
    return nsq3(n) < GjkInfo::degeneracy_sq() * nsq3(e1) * nsq3(e2);
}
__device__ __forceinline__ bool is_coplanar(f3 v1, f3 v2, f3 v3, f3 v4) {
    const f3 e1 = normalize3(v2 - v1), e2 = normalize3(v3 - v1);
    const f3 n = cross3(e1, e2);
    const f3 diff = v4 - v1;
    const float nd = dot3(n, diff);
    return nd * nd < GjkInfo::degeneracy_sq() * nsq3(n) * nsq3(diff);
}

// gjk.py:1471-1510
__device__ __forceinline__ bool simplex_degenerate(const Simplex& s, f3 mk) {
    bool deg = false;
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (i < s.n && nsq3(s.v[i].mk - mk) < GjkInfo::degeneracy_sq()) deg = true;
    if (deg) return true;
    if (s.n == 2) return is_colinear(s.v[0].mk, s.v[1].mk, mk);
    if (s.n == 3) return is_coplanar(s.v[0].mk, s.v[1].mk, s.v[2].mk, mk);
    return false;
}

__device__ __forceinline__ bool simplex_valid(const Simplex& s, const SVert& c) {
    return !simplex_duplicate(s, c.id1, c.id2) && !simplex_degenerate(s, c.mk);
}

// gjk.py:1745-1857 func_safe_gjk_support: the direction is perturbed by EPS steps (signs -(1 - 2 (j & bit)), bit =
// 1, 2, 4, as the reference writes them) until the support is unique and keeps the simplex valid
template <class Sup>
__device__ __forceinline__ SVert safe_gjk_support(const DModel& m, const Simplex& s, const GjkPair& p, f3 dir,
                                                  const Sup& sup) {
    const float EPS = m.eps;
    SVert c;
    c.o1 = c.o2 = zero3();
    c.id1 = c.id2 = -1;
    c.mk = c.o1 - c.o2;
#pragma clang loop unroll(disable)
    for (int i = 0; i < 9; i++) {
        f3 nd = dir;
        if (i > 0) {
            const int j = i - 1;
            nd.x += -(1.f - 2.f * (float)(j & 1)) * EPS;
            nd.y += -(1.f - 2.f * (float)(j & 2)) * EPS;
            nd.z += -(1.f - 2.f * (float)(j & 4)) * EPS;
        }
        nd = nd * (2.f - dot3(nd, dir));
        const int nsup = count_support_geom(nd, p.ia, p.qa) * count_support_geom(-nd, p.ib, p.qb);
        if (i > 0 && nsup > 1) continue;
        c = sup(nd);
        if (i == 0) {
            if (nsup > 1) continue;
            break;
        }
        if (i == 8) break;
        if (simplex_valid(s, c)) break;
    }
    return c;
}

// gjk.py:1711-1742 func_safe_gjk_triangle_info
__device__ __forceinline__ void safe_triangle_info(f3 v1, f3 v2, f3 v3, f3 ap, f3& n, float& sd) {
    n = normalize3(cross3(v3 - v1, v2 - v1));
    if (dot3(n, ap - v1) > 0.f) n = -n;
    sd = dot3(n, v1);
}

// ------------------------------------------------------------------------------------------------ geometry
// epa.py:1388-1432 func_plane_normal
__device__ __forceinline__ int plane_normal(f3 v1, f3 v2, f3 v3, f3& normal) {
    normal = zero3();
    const f3 d21 = v2 - v1, d31 = v3 - v1, d32 = v3 - v2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const f3 n = i == 0 ? cross3(d32, d21) : (i == 1 ? cross3(d21, d31) : cross3(d31, d32));
        const float nn = norm3(n);
        if (nn == 0.f) return kRcFail;
        if (nn > GjkInfo::fmin_()) {
            normal = normalize3(n);
            return kRcOk;
        }
    }
    return kRcFail;
}

// gjk_utils.py:189-235 func_project_origin_to_plane
__device__ __forceinline__ int project_origin_to_plane(f3 v1, f3 v2, f3 v3, f3& point) {
    point = zero3();
    const f3 d21 = v2 - v1, d31 = v3 - v1, d32 = v3 - v2;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f3 n, v;
        if (i == 0) n = cross3(d32, d21), v = v2;
        else if (i == 1) n = cross3(d21, d31), v = v1;
        else n = cross3(d31, d32), v = v3;
        const float nv = dot3(n, v), nn = nsq3(n);
        if (nn == 0.f) return kRcFail;
        if (nn > GjkInfo::fmin_()) {
            point = n * (nv / nn);
            return kRcOk;
        }
        if (i == 2) {
            if (nn < GjkInfo::fmin_()) return kRcFail;
            point = n * (nv / nn);
            return kRcOk;
        }
    }
    return kRcOk;
}

// component i of v (a select, so that v stays in registers)
__device__ __forceinline__ float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// gjk_utils.py:48-107 func_triangle_affine_coords
__device__ __forceinline__ f3 triangle_affine_coords(f3 p, f3 t1, f3 t2, f3 t3) {
    float ms[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        if (i == 1) {
            const int t = i1;
            i1 = i2;
            i2 = t;
        }
        ms[i] = t2[i1] * t3[i2] - t2[i2] * t3[i1] - t1[i1] * t3[i2] + t1[i2] * t3[i1] + t1[i1] * t2[i2] -
                t1[i2] * t2[i1];
    }
    float m_max = 0.f;
    int ix = 0, iy = 0;
    bool found = false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float a = fabsf(ms[i]), b = fabsf(ms[(i + 1) % 3]), c = fabsf(ms[(i + 2) % 3]);
        if (!found && a >= b && a >= c) {
            found = true;
            m_max = ms[i];
            ix = (i + 1) % 3;
            iy = (i + 2) % 3;
            if (i == 1) {
                const int t = ix;
                ix = iy;
                iy = t;
            }
        }
    }
    float cs[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        f3 a = t2, b = t3;
        if (i == 1) a = t3, b = t1;
        else if (i == 2) a = t1, b = t2;
        const float px = comp(p, ix), py = comp(p, iy), ax = comp(a, ix), ay = comp(a, iy), bx = comp(b, ix),
                    by = comp(b, iy);
        cs[i] = px * ay + py * bx + ax * by - px * by - py * ax - bx * ay;
    }
    return mk3(cs[0] / m_max, cs[1] / m_max, cs[2] / m_max);
}

// --------------------------------------------------------------------------------------- polytope storage
// Face record: vertices (v1, v2, v3), adjacent faces (a1, a2, a3) across edges (v1 v2), (v2 v3), (v3 v1), the face's
// position in the candidate map (-1: not a candidate, -2: deleted), the outward normal and the squared distance of the
// face's plane from the origin.


// Words of MemStore's slice of the work field (each 4 bytes).
constexpr int kMemVertWords = 11;  // mk xyz, o1 xyz, o2 xyz, id1, id2
constexpr int kMemFaceWords = 12;  // v1 v2 v3 a1 a2 a3 map nx ny nz d2 (pad)
constexpr int kGjkMemWords = kEpaVerts * kMemVertWords + kEpaFaces * kMemFaceWords + kEpaFaces + 2 * kEpaHorizon +
                             2 * kEpaStack;

// Sequential scans over a polytope store S (one lane per detection).
// epa.py:989-1001: the candidate face nearest the origin (first in map order among equal distances), if its squared
// distance is below lower2 (nearest is returned unchanged otherwise). The scan runs over the face records (a face is a
// candidate when its map position is set) rather than through the map, so that the loads of different faces are
// independent; the winner is the same: the smallest distance, and among equal distances the smallest map position.
template <class S>
__device__ __forceinline__ int nearest_face_scan(const S& st, int n_map, int n_faces, int nearest, float& lower2) {
    const float lim = lower2;
    float bd = lim;
    int bi = 0x7fffffff, bf = -1;
    auto take = [&](int f, int mi, float d2) {
        if (mi >= 0 && mi < n_map && d2 < lim && (bf < 0 || d2 < bd || (d2 == bd && mi < bi))) {
            bd = d2;
            bi = mi;
            bf = f;
        }
    };
    int f = 0;
    for (; f + 4 <= n_faces; f += 4) {
        int mi[4];
        float d2[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            mi[u] = st.fmapidx(f + u);
            d2[u] = st.fdist2(f + u);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) take(f + u, mi[u], d2[u]);
    }
    for (; f < n_faces; f++) take(f, st.fmapidx(f), st.fdist2(f));
    if (bf < 0) return nearest;
    lower2 = bd;
    return bf;
}
// epa.py:1335-1350: the orientation of the polytope's vertex farthest from the face's plane (other than its own)
template <class S>
__device__ __forceinline__ float max_orient_seq(const S& st, int nverts, int v1, int v2, int v3, f3 normal, f3 fc,
                                                float init) {
    float mo = init, ma = fabsf(init);
#pragma unroll 4
    for (int iv = 0; iv < nverts; iv++) {
        if (iv != v1 && iv != v2 && iv != v3) {
            const float o = dot3(normal, st.vmk(iv) - fc);
            if (fabsf(o) > ma) {
                ma = fabsf(o);
                mo = o;
            }
        }
    }
    return mo;
}
// epa.py:1036-1046: another vertex with the same pair of support ids
template <class S>
__device__ __forceinline__ bool repeated_seq(const S& st, int nverts, int wi) {
    const int a = st.vid1(wi), b = st.vid2(wi);
    bool rep = false;
#pragma unroll 4
    for (int i = 0; i < nverts; i++)
        if (i != wi && st.vid1(i) == a && st.vid2(i) == b) rep = true;
    return rep;
}

// The polytope in memory (one lane per detection).
struct MemStore {
    static constexpr int kTeam = 1;
    static constexpr bool kAbort = false;
    static constexpr int kFaces = kEpaFaces, kVerts = kEpaVerts, kHorizon = kEpaHorizon, kStack = kEpaStack;
    float* w;  // kGjkMemWords words
    __device__ __forceinline__ float* vrec(int i) const { return w + i * kMemVertWords; }
    __device__ __forceinline__ float* frec(int f) const { return w + kEpaVerts * kMemVertWords + f * kMemFaceWords; }
    __device__ __forceinline__ int* fmap() const { return (int*)(w + kEpaVerts * kMemVertWords + kEpaFaces * kMemFaceWords); }
    __device__ __forceinline__ int* hor() const { return fmap() + kEpaFaces; }
    __device__ __forceinline__ int* stk() const { return hor() + 2 * kEpaHorizon; }

    __device__ __forceinline__ void set_vert(int i, const SVert& s) const {
        float* r = vrec(i);
        r[0] = s.mk.x; r[1] = s.mk.y; r[2] = s.mk.z;
        r[3] = s.o1.x; r[4] = s.o1.y; r[5] = s.o1.z;
        r[6] = s.o2.x; r[7] = s.o2.y; r[8] = s.o2.z;
        ((int*)r)[9] = s.id1;
        ((int*)r)[10] = s.id2;
    }
    __device__ __forceinline__ f3 vmk(int i) const { const float* r = vrec(i); return mk3(r[0], r[1], r[2]); }
    __device__ __forceinline__ f3 vo1(int i) const { const float* r = vrec(i); return mk3(r[3], r[4], r[5]); }
    __device__ __forceinline__ f3 vo2(int i) const { const float* r = vrec(i); return mk3(r[6], r[7], r[8]); }
    __device__ __forceinline__ int vid1(int i) const { return ((const int*)vrec(i))[9]; }
    __device__ __forceinline__ int vid2(int i) const { return ((const int*)vrec(i))[10]; }

    __device__ __forceinline__ void set_face(int f, int v1, int v2, int v3, int a1, int a2, int a3) const {
        int* r = (int*)frec(f);
        r[0] = v1; r[1] = v2; r[2] = v3;
        r[3] = a1; r[4] = a2; r[5] = a3;
    }
    __device__ __forceinline__ int fvert(int f, int k) const { return ((const int*)frec(f))[k]; }
    __device__ __forceinline__ void fverts(int f, int& v1, int& v2, int& v3) const {
        const int* r = (const int*)frec(f);
        v1 = r[0]; v2 = r[1]; v3 = r[2];
    }
    __device__ __forceinline__ int fadj(int f, int k) const { return ((const int*)frec(f))[3 + k]; }
    __device__ __forceinline__ void set_fadj(int f, int k, int a) const { ((int*)frec(f))[3 + k] = a; }
    __device__ __forceinline__ int fmapidx(int f) const { return ((const int*)frec(f))[6]; }
    __device__ __forceinline__ void set_fmapidx(int f, int v) const { ((int*)frec(f))[6] = v; }
    __device__ __forceinline__ f3 fnormal(int f) const { const float* r = frec(f); return mk3(r[7], r[8], r[9]); }
    __device__ __forceinline__ float fdist2(int f) const { return frec(f)[10]; }
    __device__ __forceinline__ void set_fgeom(int f, f3 n, float d2) const {
        float* r = frec(f);
        r[7] = n.x; r[8] = n.y; r[9] = n.z; r[10] = d2;
    }
    __device__ __forceinline__ int map_at(int i) const { return fmap()[i]; }
    __device__ __forceinline__ void set_map_at(int i, int f) const { fmap()[i] = f; }
    __device__ __forceinline__ void set_horizon(int i, int f, int e) const { hor()[2 * i] = f; hor()[2 * i + 1] = e; }
    __device__ __forceinline__ void horizon(int i, int& f, int& e) const { f = hor()[2 * i]; e = hor()[2 * i + 1]; }
    __device__ __forceinline__ void push(int i, int f, int e) const { stk()[2 * i] = f; stk()[2 * i + 1] = e; }
    __device__ __forceinline__ void peek(int i, int& f, int& e) const { f = stk()[2 * i]; e = stk()[2 * i + 1]; }

    // epa.py:989-1001: the candidate face nearest the origin (first in map order among equal distances), if its
    // squared distance is below lower2; nearest is returned unchanged otherwise
    __device__ __forceinline__ int nearest_face(int n_map, int n_faces, int nearest, float& lower2) const {
        return nearest_face_scan(*this, n_map, n_faces, nearest, lower2);
    }
    __device__ __forceinline__ float max_orient(int nverts, int v1, int v2, int v3, f3 normal, f3 fc, float init) const {
        return max_orient_seq(*this, nverts, v1, v2, v3, normal, fc, init);
    }
    __device__ __forceinline__ bool repeated(int nverts, int wi) const { return repeated_seq(*this, nverts, wi); }
};

// ------------------------------------------------------------------------------------- wavefront reductions
// Reductions over the 64 lanes with DPP: a scan within each row of 16 lanes (row_shr 1, 2, 4, 8), then the rows
// combined through row_bcast:15 and row_bcast:31; lane 63 ends with the result. Lanes whose DPP source lies outside
// their row keep their own value (bound_ctrl off, old = the value itself), so they combine the value with itself.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), kCtrl, kRowMask, 0xf, false));
}
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ int dpp_i(int x) {
    return __builtin_amdgcn_update_dpp(x, x, kCtrl, kRowMask, 0xf, false);
}
__device__ __forceinline__ float wave_max_f(float x) {
    x = fmaxf(x, dpp_f<0x111>(x));
    x = fmaxf(x, dpp_f<0x112>(x));
    x = fmaxf(x, dpp_f<0x114>(x));
    x = fmaxf(x, dpp_f<0x118>(x));
    x = fmaxf(x, dpp_f<0x142, 0xa>(x));
    x = fmaxf(x, dpp_f<0x143, 0xc>(x));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), kWave - 1));
}
__device__ __forceinline__ float wave_min_f(float x) {
    x = fminf(x, dpp_f<0x111>(x));
    x = fminf(x, dpp_f<0x112>(x));
    x = fminf(x, dpp_f<0x114>(x));
    x = fminf(x, dpp_f<0x118>(x));
    x = fminf(x, dpp_f<0x142, 0xa>(x));
    x = fminf(x, dpp_f<0x143, 0xc>(x));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), kWave - 1));
}
__device__ __forceinline__ int wave_min_i(int x) {
    x = min(x, dpp_i<0x111>(x));
    x = min(x, dpp_i<0x112>(x));
    x = min(x, dpp_i<0x114>(x));
    x = min(x, dpp_i<0x118>(x));
    x = min(x, dpp_i<0x142, 0xa>(x));
    x = min(x, dpp_i<0x143, 0xc>(x));
    return __builtin_amdgcn_readlane(x, kWave - 1);
}

// The polytope in the wavefront's LDS slice, at capacities that cover the detections met in practice (the largest
// Franka polytopes seen have under 90 faces and 25 vertices): kLdsFaces faces, kLdsVerts vertices, kLdsHorizon horizon
// edges and kLdsStack stack entries. When a detection would outgrow one of them it is abandoned and run again from the
// start with MemStore at the reference's capacities, so the result is always the reference algorithm's. Every lane
// runs the sequential parts with the same values and reads an entry at a uniform address (an LDS broadcast); uniform
// writes are made by every lane with the same value. The scans run one entry per lane and reduce with DPP.
constexpr int kLdsFaces = 128, kLdsVerts = 32, kLdsHorizon = 32, kLdsStack = 256;
constexpr int kLdsVertWords = 12;  // mk xyz, o1 xyz, o2 xyz, id1, id2, pad (16-byte aligned)
constexpr int kLdsFaceWords = 8;   // vertices and map position (bytes), adjacent faces (halfwords), normal, d2
constexpr int kGjkLdsWords = kLdsVerts * kLdsVertWords + kLdsFaces * kLdsFaceWords + kLdsFaces / 4 + kLdsHorizon / 2 +
                             kLdsStack / 2;
static_assert(kLdsVerts <= kWave && kLdsFaces <= 2 * kWave && kLdsFaces <= 254, "LDS polytope packing");

template <bool kWaveOps>
struct CompactStore {
    static constexpr int kTeam = kWaveOps ? kWave : 1;
    static constexpr bool kAbort = true;
    static constexpr int kFaces = kLdsFaces, kVerts = kLdsVerts, kHorizon = kLdsHorizon, kStack = kLdsStack;
    float* w;  // kGjkLdsWords words, 16-byte aligned
    int lane;

    __device__ __forceinline__ float* vrec(int i) const { return w + i * kLdsVertWords; }
    __device__ __forceinline__ float* frec(int f) const { return w + kLdsVerts * kLdsVertWords + f * kLdsFaceWords; }
    __device__ __forceinline__ uint8_t* fmap() const {
        return (uint8_t*)(w + kLdsVerts * kLdsVertWords + kLdsFaces * kLdsFaceWords);
    }
    __device__ __forceinline__ uint16_t* hor() const { return (uint16_t*)(fmap() + kLdsFaces); }
    __device__ __forceinline__ uint16_t* stk() const { return hor() + kLdsHorizon; }

    __device__ __forceinline__ void set_vert(int i, const SVert& s) const {
        float4* r = (float4*)vrec(i);
        r[0] = make_float4(s.mk.x, s.mk.y, s.mk.z, s.o1.x);
        r[1] = make_float4(s.o1.y, s.o1.z, s.o2.x, s.o2.y);
        r[2] = make_float4(s.o2.z, __int_as_float(s.id1), __int_as_float(s.id2), 0.f);
    }
    __device__ __forceinline__ f3 vmk(int i) const { const float* r = vrec(i); return mk3(r[0], r[1], r[2]); }
    __device__ __forceinline__ f3 vo1(int i) const { const float* r = vrec(i); return mk3(r[3], r[4], r[5]); }
    __device__ __forceinline__ f3 vo2(int i) const { const float* r = vrec(i); return mk3(r[6], r[7], r[8]); }
    __device__ __forceinline__ int vid1(int i) const { return __float_as_int(vrec(i)[9]); }
    __device__ __forceinline__ int vid2(int i) const { return __float_as_int(vrec(i)[10]); }

    // face record: bytes 0-2 the vertices, byte 3 the map position + 2, bytes 4-9 the adjacent faces (16 bits each),
    // words 3-5 the normal, word 6 the squared distance; every field has its own byte, halfword or word, so that a
    // field is written without reading the record back
    __device__ __forceinline__ int* fw(int f) const { return (int*)frec(f); }
    __device__ __forceinline__ uint8_t* fb(int f) const { return (uint8_t*)frec(f); }
    __device__ __forceinline__ uint16_t* fa(int f) const { return (uint16_t*)(frec(f) + 1); }
    // a new face starts outside the candidate map (map position -1, what a successful attach sets; a failed attach
    // ends the EPA before the position is read)
    __device__ __forceinline__ void set_face(int f, int v1, int v2, int v3, int a1, int a2, int a3) const {
        int* r = fw(f);
        r[0] = v1 | (v2 << 8) | (v3 << 16) | (1 << 24);
        r[1] = a1 | (a2 << 16);
        r[2] = a3;
    }
    __device__ __forceinline__ void fverts(int f, int& v1, int& v2, int& v3) const {
        const int p = fw(f)[0];
        v1 = p & 0xff;
        v2 = (p >> 8) & 0xff;
        v3 = (p >> 16) & 0xff;
    }
    __device__ __forceinline__ int fvert(int f, int k) const { return fb(f)[k]; }
    __device__ __forceinline__ int fadj(int f, int k) const { return fa(f)[k]; }
    __device__ __forceinline__ void set_fadj(int f, int k, int a) const { fa(f)[k] = (uint16_t)a; }
    __device__ __forceinline__ int fmapidx(int f) const { return (int)fb(f)[3] - 2; }
    __device__ __forceinline__ void set_fmapidx(int f, int v) const { fb(f)[3] = (uint8_t)(v + 2); }
    __device__ __forceinline__ f3 fnormal(int f) const { const float* r = frec(f); return mk3(r[3], r[4], r[5]); }
    __device__ __forceinline__ float fdist2(int f) const { return frec(f)[6]; }
    __device__ __forceinline__ void set_fgeom(int f, f3 n, float d2) const {
        float* r = frec(f);
        r[3] = n.x; r[4] = n.y; r[5] = n.z; r[6] = d2;
    }
    __device__ __forceinline__ int map_at(int i) const { return fmap()[i]; }
    __device__ __forceinline__ void set_map_at(int i, int f) const { fmap()[i] = (uint8_t)f; }
    __device__ __forceinline__ void set_horizon(int i, int f, int e) const { hor()[i] = (uint16_t)(f | (e << 8)); }
    __device__ __forceinline__ void horizon(int i, int& f, int& e) const {
        const int p = hor()[i];
        f = p & 0xff;
        e = p >> 8;
    }
    __device__ __forceinline__ void push(int i, int f, int e) const { stk()[i] = (uint16_t)(f | (e << 8)); }
    __device__ __forceinline__ void peek(int i, int& f, int& e) const {
        const int p = stk()[i];
        f = p & 0xff;
        e = p >> 8;
    }

    // smallest (d2, map position) over the candidate faces with d2 < lower2 (the sequential scan's first minimum in
    // map order): each lane takes faces lane and lane + 64, then two reductions (the distance, then the position
    // among the faces at that distance)
    __device__ __forceinline__ int nearest_face(int n_map, int n_faces, int nearest, float& lower2) const {
        if constexpr (!kWaveOps) return nearest_face_seq(n_map, n_faces, nearest, lower2);
        const float inf = __int_as_float(0x7f800000);
        float bd = inf;
        int bi = 0x7fffffff, bf = -1;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int f = k * kWave + lane;
            if (f < n_faces) {
                const int mi = fmapidx(f);
                const float d2 = fdist2(f);
                if (mi >= 0 && mi < n_map && d2 < lower2 && (d2 < bd || (d2 == bd && mi < bi))) {
                    bd = d2;
                    bi = mi;
                    bf = f;
                }
            }
        }
        const float md = wave_min_f(bd);
        if (!(md < lower2)) return nearest;
        const int mpos = wave_min_i(bd == md ? bi : 0x7fffffff);
        const uint64_t win = wave_ballot(bd == md && bi == mpos);
        lower2 = md;
        return rl_i(bf, __builtin_ctzll(win));
    }
    // epa.py:1335-1350 with one vertex per lane: the first vertex (other than the face's) farthest from the plane
    __device__ __forceinline__ float max_orient(int nverts, int v1, int v2, int v3, f3 normal, f3 fc, float init) const {
        if constexpr (!kWaveOps) return max_orient_seq(*this, nverts, v1, v2, v3, normal, fc, init);
        const float ma = fabsf(init);
        float key = -1.f, o = 0.f;
        if (lane < nverts && lane != v1 && lane != v2 && lane != v3) {
            o = dot3(normal, vmk(lane) - fc);
            if (fabsf(o) > ma) key = fabsf(o);
        }
        const float mk = wave_max_f(key);
        if (!(mk >= 0.f)) return init;
        return rl_f(o, __builtin_ctzll(wave_ballot(key == mk)));
    }
    __device__ __forceinline__ bool repeated(int nverts, int wi) const {
        if constexpr (!kWaveOps) return repeated_seq(*this, nverts, wi);
        const int a = vid1(wi), b = vid2(wi);
        return wave_ballot(lane < nverts && lane != wi && vid1(lane) == a && vid2(lane) == b) != 0;
    }
    // one lane alone: the face records scanned with independent loads (as in MemStore)
    __device__ __forceinline__ int nearest_face_seq(int n_map, int n_faces, int nearest, float& lower2) const {
        return nearest_face_scan(*this, n_map, n_faces, nearest, lower2);
    }
};
using LdsStore = CompactStore<true>;    // one detection per wavefront, the polytope in the wavefront's LDS
using LaneStore = CompactStore<false>;  // one detection per lane, the polytope in the lane's slot of the work field

// -------------------------------------------------------------------------------------------- the detection
// The detection's state; the model and the pair are passed to every step (a struct holding references to them would
// keep the state in scratch memory).
template <class Store>
struct GjkEpa {
    Store S;
    int lane;
    bool overflow;  // a capacity of the store was reached (abortable stores only): the detection must run again
    static constexpr int team = Store::kTeam;

    // both support points along d (lane-parallel on a wavefront)
    __device__ __forceinline__ SVert sup_(const DModel& m, const GjkPair& p, f3 d) const {
        if constexpr (team == kWave) return support_pair_wave(m, p, d, lane);
        else return support_pair_ids(m, p, d);
    }
    Simplex s;
    int last_searched;
    int nverts, nfaces, n_map, h_nedges;
    float distance;
    int n_witness;
    f3 w1, w2;

    // gjk.py:1551-1657 func_search_valid_simplex_vertex
    __device__ __forceinline__ bool search_valid_vertex(const DModel& m, const GjkPair& p, SVert& out) {
        SVert c;
        c.o1 = c.o2 = c.mk = zero3();
        c.id1 = c.id2 = -1;
        if (is_discrete(p.ia) && is_discrete(p.ib)) {
            const int na = p.ia.vend - p.ia.vstart;
            const int nb = p.ib.vend - p.ib.vstart;
            const int ncase = na * nb;
            for (int k0 = 0; k0 < ncase; k0 += team) {
                const int k = k0 + (team > 1 ? lane : 0);
                bool ok = false;
                int mm = 0;
                if (k < ncase) {
                    mm = (k + last_searched) % ncase;
                    const int i = mm / nb, j = mm % nb;
                    c.id1 = p.ia.vstart + i;
                    c.id2 = p.ib.vstart + j;
                    c.o1 = discrete_vertex(m, p.ia, p.pa, p.qa, i);
                    c.o2 = discrete_vertex(m, p.ib, p.pb, p.qb, j);
                    c.mk = c.o1 - c.o2;
                    ok = simplex_valid(s, c);
                }
                if constexpr (team == 1) {
                    if (ok) {
                        last_searched = (mm + 1) % ncase;
                        out = c;
                        return true;
                    }
                } else {
                    const uint64_t bal = wave_ballot(ok);
                    if (bal) {
                        const int wl = __builtin_ctzll(bal);
                        out.o1 = mk3(rl_f(c.o1.x, wl), rl_f(c.o1.y, wl), rl_f(c.o1.z, wl));
                        out.o2 = mk3(rl_f(c.o2.x, wl), rl_f(c.o2.y, wl), rl_f(c.o2.z, wl));
                        out.mk = mk3(rl_f(c.mk.x, wl), rl_f(c.mk.y, wl), rl_f(c.mk.z, wl));
                        out.id1 = rl_i(c.id1, wl);
                        out.id2 = rl_i(c.id2, wl);
                        last_searched = (rl_i(mm, wl) + 1) % ncase;
                        return true;
                    }
                }
            }
            out = c;
            return false;
        }
        if (s.n == 3) {
            const f3 dir = normalize3(cross3(s.v[2].mk - s.v[0].mk, s.v[1].mk - s.v[0].mk));
            for (int i = 0; i < 2; i++) {
                c = sup_(m, p, i == 0 ? dir : -dir);
                if (simplex_valid(s, c)) {
                    out = c;
                    return true;
                }
            }
        }
        out = c;
        return false;
    }

    // gjk.py:1209-1425 func_safe_gjk
    __device__ __forceinline__ int safe_gjk(const DModel& m, const GjkPair& p) {
        bool init_ok = true;
        s.n = 0;
#pragma clang loop unroll(disable)
        for (int i = 0; i < 4; i++) {
            const float sg = 1.f - 2.f * (float)(i % 2);
            const f3 dir = i < 2 ? mk3(0.f, 0.f, sg) : mk3(0.f, sg, 0.f);  // dir[2 - i / 2] = sg
            SVert c = safe_gjk_support(m, s, p, dir, [&](f3 d) { return sup_(m, p, d); });
            if (!simplex_valid(s, c)) {
                if (!search_valid_vertex(m, p, c)) {
                    init_ok = false;
                    break;
                }
            }
            s.set(i, c);
            s.n++;
        }
        int flag = kGjkSeparated;
        if (init_ok) {
#pragma clang loop unroll(disable)
            for (int it = 0; it < kGjkIters; it++) {
                f3 fn[4];
                float fd[4];
                // face j is the triangle opposite vertex j
                safe_triangle_info(s.v[2].mk, s.v[1].mk, s.v[3].mk, s.v[0].mk, fn[0], fd[0]);
                safe_triangle_info(s.v[0].mk, s.v[2].mk, s.v[3].mk, s.v[1].mk, fn[1], fd[1]);
                safe_triangle_info(s.v[1].mk, s.v[0].mk, s.v[3].mk, s.v[2].mk, fn[2], fd[2]);
                safe_triangle_info(s.v[0].mk, s.v[1].mk, s.v[2].mk, s.v[3].mk, fn[3], fd[3]);
                int mi = 0;
                f3 mn = fn[0];
                float md = fd[0];
#pragma unroll
                for (int j = 1; j < 4; j++)
                    if (fd[j] < md) {
                        mi = j;
                        mn = fn[j];
                        md = fd[j];
                    }
                if (md >= 0.f) {
                    flag = kGjkIntersect;
                    break;
                }
                s.n = 3;
                s.set(mi, s.v[3]);  // vertex 3 takes the place of the dropped apex (a no-op when that is vertex 3)
                const SVert c = safe_gjk_support(m, s, p, mn, [&](f3 d) { return sup_(m, p, d); });
                if (simplex_duplicate(s, c.id1, c.id2)) {
                    flag = kGjkSeparated;
                    break;
                }
                if (simplex_degenerate(s, c.mk)) {
                    flag = kGjkNumError;
                    break;
                }
                if (dot3(c.mk, mn) < 0.f) {
                    flag = kGjkSeparated;
                    break;
                }
                s.v[3] = c;
                s.n = 4;
            }
        }
        if (flag == kGjkIntersect) {
            distance = 0.f;
        } else {
            flag = kGjkSeparated;
            distance = GjkInfo::fmax_();
        }
        return flag;
    }

    // epa.py:407-431
    __device__ __forceinline__ int insert_vertex(const SVert& c) {
        const int n = nverts;
        S.set_vert(n, c);
        nverts++;
        return n;
    }

    // epa.py:1303-1385 func_safe_attach_face_to_polytope
    __device__ __forceinline__ int attach_face(int v1, int v2, int v3, int a1, int a2, int a3) {
        const int n = nfaces;
        S.set_face(n, v1, v2, v3, a1, a2, a3);
        nfaces++;
        const f3 p1 = S.vmk(v1), p2 = S.vmk(v2), p3 = S.vmk(v3);
        f3 normal;
        const int flag = plane_normal(p3, p2, p1, normal);
        if (flag == kRcOk) {
            const f3 fc = (p1 + p2 + p3) / 3.f;
            const float mo = S.max_orient(nverts, v1, v2, v3, normal, fc, -dot3(normal, fc));
            if (mo > 0.f) normal = -normal;
            float min_d2 = GjkInfo::fmax_();
            const f3 pv[3] = {p1, p2, p3};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float dd = dot3(normal, pv[k]);
                const float d2 = dd * dd;
                if (d2 < min_d2) min_d2 = d2;
            }
            S.set_fgeom(n, normal, min_d2);
            S.set_fmapidx(n, -1);
        }
        return flag;
    }

    // epa.py:1250-1300 func_safe_epa_init: the tetrahedron's four faces
    __device__ __forceinline__ void epa_init() {
#pragma unroll
        for (int i = 0; i < 4; i++) insert_vertex(s.v[i]);
        attach_face(0, 1, 2, 1, 3, 2);
        attach_face(0, 3, 1, 2, 3, 0);
        attach_face(0, 2, 3, 0, 3, 1);
        attach_face(3, 2, 1, 2, 0, 1);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            S.set_map_at(i, i);
            S.set_fmapidx(i, i);
        }
        n_map = 4;
    }

    // epa.py:383-404
    __device__ __forceinline__ void delete_face(int f) {
        const int mi = S.fmapidx(f);
        if (mi >= 0) {
            const int last = S.map_at(n_map - 1);
            S.set_map_at(mi, last);
            S.set_fmapidx(last, mi);
            n_map--;
        }
        S.set_fmapidx(f, -2);
    }

    // epa.py:361-380
    __device__ __forceinline__ int edge_idx(int f, int v) const {
        int a, b, c;
        S.fverts(f, a, b, c);
        return a == v ? 0 : (b == v ? 1 : 2);
    }

    // epa.py:273-340 func_epa_horizon: depth-first walk of the faces visible from w
    __device__ __forceinline__ int horizon(f3 w, int nearest) {
        S.push(0, nearest, 0);
        int top = 1;
        bool first = true;
#pragma clang loop unroll(disable)
        while (top > 0) {
            int f, e;
            S.peek(top - 1, f, e);
            int fv0, fv1, fv2;
            S.fverts(f, fv0, fv1, fv2);
            const f3 v = S.vmk(fv0);
            top--;
            const bool deleted = S.fmapidx(f) == -2;
            if (!first && deleted) continue;
            const bool visible = dot3(S.fnormal(f), w - v) > GjkInfo::fmin_();
            if (visible || first) {
                delete_face(f);
                for (int k = first ? 0 : 1; k < 3; k++) {
                    const int e2 = (e + k) % 3;
                    const int adj = S.fadj(f, e2);
                    if (S.fmapidx(adj) != -2) {
                        const int sv = e2 == 2 ? fv0 : (e2 == 1 ? fv2 : fv1);  // f_verts[f][(e2 + 1) % 3]
                        if (top >= Store::kStack) {
                            if constexpr (Store::kAbort) overflow = true;
                            return kRcFail;
                        }
                        S.push(top, adj, edge_idx(adj, sv));
                        top++;
                    }
                }
            } else {
                if (h_nedges >= Store::kHorizon) {
                    if constexpr (Store::kAbort) overflow = true;
                    return kRcFail;
                }
                S.set_horizon(h_nedges, f, e);
                h_nedges++;
            }
            first = false;
        }
        return kRcOk;
    }

    // epa.py:1183-1247 func_safe_epa_witness
    __device__ __forceinline__ int witness(int f) {
        int i1, i2, i3;
        S.fverts(f, i1, i2, i3);
        const f3 v1 = S.vmk(i1), v2 = S.vmk(i2), v3 = S.vmk(i3);
        f3 proj;
        project_origin_to_plane(v1, v2, v3, proj);
        const f3 lam = triangle_affine_coords(proj, v1, v2, v3);
        const f3 rp = v1 * lam.x + v2 * lam.y + v3 * lam.z;
        const float err = norm3(proj - rp);
        const float e12 = nsq3(v1 - v2), e23 = nsq3(v2 - v3), e31 = nsq3(v3 - v1);
        const float mx = fmaxf(fmaxf(e12, e23), fmaxf(e31, GjkInfo::fmin_sq()));
        const float rel = err * (1.f / sqrtf(mx));
        if (rel > GjkInfo::max_rel_reproj() && err > GjkInfo::max_abs_reproj()) return kRcFail;
        w1 = S.vo1(i1) * lam.x + S.vo1(i2) * lam.y + S.vo1(i3) * lam.z;
        w2 = S.vo2(i1) * lam.x + S.vo2(i2) * lam.y + S.vo2(i3) * lam.z;
        return kRcOk;
    }

    // epa.py:969-1180 func_safe_epa
    __device__ __forceinline__ int safe_epa(const DModel& m, const GjkPair& p) {
        float upper = GjkInfo::fmax_(), upper2 = GjkInfo::fmax_sq(), lower = 0.f;
        float tol = GjkInfo::tol();
        const float EPS = m.eps;
        int nearest = -1, prev = -1;
        const bool discrete = is_discrete(p.ia) && is_discrete(p.ib);
        if (discrete) tol = EPS;
#pragma clang loop unroll(disable)
        for (int k = 0; k < kEpaIters; k++) {
            prev = nearest;
            float lower2 = GjkInfo::fmax_sq();
            nearest = S.nearest_face(n_map, nfaces, nearest, lower2);
            if (lower2 > upper2 || nearest == -1) {
                nearest = prev;
                break;
            }
            lower = sqrtf(lower2);
            const f3 dir = S.fnormal(nearest);
            const SVert c = sup_(m, p, dir);
            const int wi = insert_vertex(c);
            const f3 w = c.mk;
            const float upk = dot3(w, dir);
            if (upk < upper) {
                upper = upk;
                upper2 = upper * upper;
            }
            if (upper - lower < tol) break;
            if (discrete && S.repeated(nverts, wi)) break;
            if (horizon(w, nearest) != kRcOk) {
                nearest = -1;
                break;
            }
            if (h_nedges < 3) {
                nearest = -1;
                break;
            }
            const int nf = nfaces, ne = h_nedges;
            if (nf + ne >= Store::kFaces) {
                if constexpr (Store::kAbort) overflow = true;
                break;
            }
            int aflag = kRcOk;
            for (int i = 0; i < ne; i++) {
                const int f0 = nf + i;
                const int f1 = nf + (i + 1) % ne;
                int hf, he;
                S.horizon(i, hf, he);
                const int hv1 = S.fvert(hf, he);
                const int hv2 = S.fvert(hf, (he + 1) % 3);
                S.set_fadj(hf, he, f0);
                const int adj0 = i > 0 ? f0 - 1 : nf + ne - 1;
                aflag = attach_face(wi, hv2, hv1, f1, hf, adj0);
                if (aflag != kRcOk) break;
                const float d2 = S.fdist2(nfaces - 1);
                if (d2 >= lower2 - EPS && d2 <= upper2 + EPS) {
                    S.set_map_at(n_map, f0);
                    S.set_fmapidx(f0, n_map);
                    n_map++;
                }
            }
            if (aflag != kRcOk) {
                nearest = -1;
                break;
            }
            h_nedges = 0;
            if (n_map == 0 || nearest == -1) {
                nearest = -1;
                break;
            }
            if (nverts >= Store::kVerts) {  // the next iteration would insert past the vertex array
                if constexpr (Store::kAbort) overflow = true;
                break;
            }
        }
        if constexpr (Store::kAbort) {
            if (overflow) return -1;
        }
        if (nearest != -1) {
            const float d2 = S.fdist2(nearest);
            if (witness(nearest) == kRcOk) {
                n_witness = 1;
                distance = -sqrtf(d2);
            } else {
                n_witness = 0;
                distance = 0.f;
                nearest = -1;
            }
        } else {
            n_witness = 0;
            distance = 0.f;
        }
        return nearest;
    }

    // gjk.py:170-446 func_gjk_contact (default path): whether the geoms collide, with the contact normal (from a to
    // b), position (midpoint of the witness points) and penetration
    __device__ __forceinline__ bool contact(const DModel& m, const GjkPair& p, f3& normal, f3& pos, float& pen) {
        last_searched = 0;
        n_witness = 0;
        overflow = false;
        const int flag = safe_gjk(m, p);
        if (flag == kGjkIntersect) {
            nverts = nfaces = n_map = h_nedges = 0;
            epa_init();
            safe_epa(m, p);
        }
        bool col = distance < 0.f;
        pen = col ? -distance : 0.f;
        int n_contacts = 0;
        if (col && n_witness > 0) {
            const f3 cp = 0.5f * (w1 + w2);
            const f3 n = w2 - w1;
            const float nl = norm3(n);
            if (!(nl < GjkInfo::fmin_())) {
                pos = cp;
                normal = n / nl;
                n_contacts = 1;
            }
        }
        if (n_contacts == 0) {
            col = false;
            pen = 0.f;
        }
        return col;
    }
};

// An environment's GJK/EPA area of the work field: one polytope at the reference's capacities (the per-stage kernels'
// and the fallback of the compact ones), then one compact polytope per lane of the fused kernel's wavefront.
constexpr int kGjkMemSlotWords = (kGjkMemWords + 3) / 4 * 4;
constexpr int kGjkWorkWords = kGjkMemSlotWords + kWave * kGjkLdsWords;
__device__ __forceinline__ float* gjk_lane_slot(float* gjk_mem, int lane) {
    return gjk_mem + kGjkMemSlotWords + (size_t)lane * kGjkLdsWords;
}

// One detection per lane of a wavefront (the lanes with want set), each with its polytope in the lane's compact slot of
// gjk_mem; a detection that outgrows its slot runs again afterwards on the whole wavefront with the polytope at the
// reference's capacities. Each wanting lane's outputs are set; the others' are left alone.
__device__ __forceinline__ void gjk_contact_lanes(const DModel& m, const GjkPair& p, bool want, int lane,
                                                  float* gjk_mem, bool& col, f3& normal, f3& pos, float& pen);

// One detection by one lane (or by every lane of a wavefront with the same inputs), the polytope in memory (w:
// kGjkMemWords words of the environment's work slice) at the reference's capacities.
__device__ __forceinline__ bool gjk_contact_mem(const DModel& m, const GjkPair& p, float* w, f3& normal, f3& pos,
                                                float& pen) {
    GjkEpa<MemStore> g;
    g.S.w = w;
    g.lane = 0;
    return g.contact(m, p, normal, pos, pen);
}

// One detection by the whole wavefront: the polytope in the wavefront's LDS (lds: kGjkLdsWords words, 16-byte
// aligned), and only when it outgrows that, or when no LDS was set aside (lds = nullptr), again from the start with
// the polytope in memory (mem: kGjkMemWords words) at the reference's capacities. Every lane passes the same pair and
// gets the same result.
__device__ __forceinline__ bool gjk_contact_wave(const DModel& m, const GjkPair& p, int lane, float* lds, float* mem,
                                                 f3& normal, f3& pos, float& pen) {
    if (lds) {
        GjkEpa<LdsStore> g;
        g.S.w = lds;
        g.S.lane = lane;
        g.lane = lane;
        const bool col = g.contact(m, p, normal, pos, pen);
        if (!g.overflow) return col;
    }
    return gjk_contact_mem(m, p, mem, normal, pos, pen);
}

__device__ __forceinline__ void gjk_contact_lanes(const DModel& m, const GjkPair& p, bool want, int lane,
                                                  float* gjk_mem, bool& col, f3& normal, f3& pos, float& pen) {
    bool over = false;
    if (want) {
        GjkEpa<LaneStore> g;
        g.S.w = gjk_lane_slot(gjk_mem, lane);
        g.S.lane = lane;
        g.lane = lane;
        col = g.contact(m, p, normal, pos, pen);
        over = g.overflow;
    }
#pragma clang loop unroll(disable)
    for (uint64_t om = wave_ballot(over); om; om &= om - 1) {
        const int l = __builtin_ctzll(om);
        GjkPair q;
        q.ga = rl_i(p.ga, l);
        q.gb = rl_i(p.gb, l);
        q.pa = rl3(p.pa, l);
        q.pb = rl3(p.pb, l);
        q.qa = rl4(p.qa, l);
        q.qb = rl4(p.qb, l);
        q.ia = ginfo(m, q.ga);
        q.ib = ginfo(m, q.gb);
        f3 qn = zero3(), qp = zero3();
        float qpen = 0.f;
        const bool qc = gjk_contact_mem(m, q, gjk_mem, qn, qp, qpen);
        if (lane == l) {
            col = qc;
            normal = qn;
            pos = qp;
            pen = qpen;
        }
    }
}

// narrowphase.py:726-734: after MPR, a deep penetration with no cached direction, or a large overlap, is detected
// again with GJK/EPA
__device__ __forceinline__ bool prefer_gjk(const DModel& m, float pen, float tolerance, bool avail) {
    return pen > tolerance && (!avail || m.mc_tolerance * pen >= m.mpr_to_gjk_overlap_ratio * tolerance);
}

}  // namespace rs
