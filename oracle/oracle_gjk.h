// ORACLE (test infrastructure only): CPU restatement of the reference's GJK/EPA contact detection on its default
// (non-MuJoCo-compatible) path: func_gjk_contact -> func_safe_gjk -> func_safe_epa_init -> func_safe_epa
// (genesis/engine/solvers/rigid/collider/gjk.py:170-446, :1209-1915; epa.py:273-434, :809-885, :969-1432;
// gjk_utils.py; gjk_support.py; support_field.py:121-395). The polygon-clipping multi-contact of multi_contact.py is
// reached only with enable_mujoco_multi_contact, which the reference hard-codes to False (gjk.py:51), so a GJK
// detection yields at most one contact (max_contacts_per_pair = 1, gjk.py:69-71).
//
// The state mirrors GJKState (utils/array_class.py:1069-1115) for one detection: a 4-vertex simplex, a polytope of up
// to 5 + epa_max_iterations vertices and polytope_max_faces = 6 * epa_max_iterations faces, the horizon data
// (6 + epa_max_iterations edges) and the horizon stack (3 * polytope_max_faces entries). Where the reference would
// write past one of these arrays (it does not check), this restatement stops the EPA as for a failed horizon.
#pragma once
#include <cmath>

#include "oracle_math.h"
#include "oracle_model.h"

namespace orc {

constexpr int GJK_MAX_ITER = 50, EPA_MAX_ITER = 50;
constexpr int EPA_MAX_FACES = 6 * EPA_MAX_ITER;
constexpr int EPA_MAX_VERTS = 5 + EPA_MAX_ITER;
constexpr int EPA_MAX_HORIZON = 6 + EPA_MAX_ITER;
constexpr int EPA_MAX_STACK = 3 * EPA_MAX_FACES;
constexpr int GJK_SEPARATED = 0, GJK_INTERSECT = 1, GJK_NUM_ERROR = 2;  // constants.py GJK_RETURN_CODE
constexpr int RC_SUCCESS = 0, RC_FAIL = 1;                               // constants.py RETURN_CODE

// gjk.py:76-106 get_gjk_info values used on this path
template <class T> struct GJKInfo {
    static T float_min() { return T(1e-15); }
    static T float_min_sq() { return T(1e-15) * T(1e-15); }
    static T float_max() { return T(1e15); }
    static T float_max_sq() { return T(1e15) * T(1e15); }
    static T tolerance() { return T(1e-6); }
    static T simplex_max_degeneracy_sq() { return T(1e-5) * T(1e-5); }
    static T max_rel_reproj() { return T(1e-4); }
    static T max_abs_reproj() { return T(1e-4); }
};

template <class T> struct GJKState {
    // simplex (simplex_vertex.*)
    int snv;
    V3<T> s_obj1[4], s_obj2[4], s_mink[4];
    int s_id1[4], s_id2[4];
    int last_searched;
    // polytope
    int nverts, nfaces, nfaces_map, h_nedges;
    V3<T> p_obj1[EPA_MAX_VERTS], p_obj2[EPA_MAX_VERTS], p_mink[EPA_MAX_VERTS];
    int p_id1[EPA_MAX_VERTS], p_id2[EPA_MAX_VERTS];
    int f_verts[EPA_MAX_FACES][3], f_adj[EPA_MAX_FACES][3], f_map_idx[EPA_MAX_FACES];
    V3<T> f_normal[EPA_MAX_FACES];
    T f_dist2[EPA_MAX_FACES];
    int faces_map[EPA_MAX_FACES];
    int hd_face[EPA_MAX_HORIZON], hd_edge[EPA_MAX_HORIZON];
    int hs_face[EPA_MAX_STACK], hs_edge[EPA_MAX_STACK];
    V3<T> horizon_w;
    // result
    T distance;
    int n_witness;
    V3<T> w1, w2;
};

// ------------------------------------------------------------------------------------------------- support

// support_field.py:139-181 _func_support_mesh, returning the vertex index within the geom (support_vid)
template <class T> V3<T> support_mesh_vid(const Model<T>& m, V3<T> d, int g, int& vid) {
    const int res = 180;
    const T pi = T(3.141592653589793);
    T theta = std::atan2(d[1], d[0]);
    T phi = std::acos(std::fmin(std::fmax(d[2], T(-1)), T(1)));
    T ii = (theta + pi) / pi / T(2) * T(res);
    T jj = phi / pi * T(res);
    T dot_max = T(-1e20);
    V3<T> v = v3(T(0), T(0), T(0));
    vid = 0;
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
            vid = m.support_vid[idx];
        }
    }
    return v;
}

// gjk_support.py:64-106 support_driver (shrink_sphere = False, not MuJoCo-compatible): world support point and the
// id of the vertex it came from (-1 for smooth geoms)
template <class T> V3<T> gjk_support_geom(const Model<T>& m, V3<T> d, int g, V3<T> pos, Q4<T> quat, int& vid) {
    int t = m.geom_type[g];
    const T* data = &m.geom_data[7 * g];
    vid = -1;
    if (t == GT_SPHERE) return pos + d * data[0];
    if (t == GT_ELLIPSOID) {
        V3<T> dl = inv_transform_by_quat(d, quat);
        T a = data[0], b = data[1], c = data[2];
        T nrm = std::sqrt(a * a * dl[0] * dl[0] + b * b * dl[1] * dl[1] + c * c * dl[2] * dl[2]);
        V3<T> sl = v3(a * a * dl[0] / nrm, b * b * dl[1] / nrm, c * c * dl[2] / nrm);
        return pos + transform_by_quat(sl, quat);
    }
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
        vid = (vl[0] > T(0) ? 1 : 0) + (vl[1] > T(0) ? 2 : 0) + (vl[2] > T(0) ? 4 : 0) + m.geom_vert_start[g];
        return transform_by_trans_quat(vl, pos, quat);
    }
    if (m.geom_support_start[g] < 0) return pos;
    V3<T> dm = transform_by_quat(d, inv_quat(quat));
    return transform_by_trans_quat(support_mesh_vid(m, dm, g, vid), pos, quat);
}

// gjk_support.py:109-176 func_support: support points on both geoms along d (a) and -d (b)
template <class T>
void gjk_support_pair(const Model<T>& m, int ga, int gb, V3<T> d, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb, V3<T>& o1,
                      V3<T>& o2, int& id1, int& id2, V3<T>& mink) {
    o1 = gjk_support_geom(m, d, ga, pa, qa, id1);
    o2 = gjk_support_geom(m, -d, gb, pb, qb, id2);
    mink = o1 - o2;
}

// support_field.py:312-395 _func_count_supports_*: a box counts 2^(zero components of the local direction); the mesh
// count compares against a dot_max that is never raised, so it is always 1
template <class T> int count_support_geom(const Model<T>& m, V3<T> d, int g, Q4<T> quat) {
    if (m.geom_type[g] == GT_BOX) {
        V3<T> db = inv_transform_by_quat(d, quat);
        int z = (db[0] == T(0)) + (db[1] == T(0)) + (db[2] == T(0));
        return 1 << z;
    }
    return 1;
}

// ------------------------------------------------------------------------------------------ simplex checks

// gjk.py:1449-1468
template <class T> bool simplex_duplicate(const GJKState<T>& s, int id1, int id2) {
    for (int i = 0; i < s.snv; i++) {
        if (id1 == -1 || s.s_id1[i] != id1) continue;
        if (id2 == -1 || s.s_id2[i] != id2) continue;
        return true;
    }
    return false;
}

// gjk.py:1513-1548
template <class T> bool is_colinear(V3<T> v1, V3<T> v2, V3<T> v3) {
    V3<T> e1 = v2 - v1, e2 = v3 - v1;
    V3<T> n = cross(e1, e2);
    return norm_sqr(n) < GJKInfo<T>::simplex_max_degeneracy_sq() * norm_sqr(e1) * norm_sqr(e2);
}
template <class T> bool is_coplanar(V3<T> v1, V3<T> v2, V3<T> v3, V3<T> v4) {
    V3<T> e1 = normalized(v2 - v1), e2 = normalized(v3 - v1);
    V3<T> n = cross(e1, e2);
    V3<T> diff = v4 - v1;
    T nd = dot(n, diff);
    return nd * nd < GJKInfo<T>::simplex_max_degeneracy_sq() * norm_sqr(n) * norm_sqr(diff);
}

// gjk.py:1471-1510
template <class T> bool simplex_degenerate(const GJKState<T>& s, V3<T> mink) {
    for (int i = 0; i < s.snv; i++)
        if (norm_sqr(s.s_mink[i] - mink) < GJKInfo<T>::simplex_max_degeneracy_sq()) return true;
    if (s.snv == 2) return is_colinear(s.s_mink[0], s.s_mink[1], mink);
    if (s.snv == 3) return is_coplanar(s.s_mink[0], s.s_mink[1], s.s_mink[2], mink);
    return false;
}

template <class T> bool simplex_valid(const GJKState<T>& s, int id1, int id2, V3<T> mink) {
    return !simplex_duplicate(s, id1, id2) && !simplex_degenerate(s, mink);
}

// gjk.py:1674-1708 func_get_discrete_geom_vertex
template <class T> V3<T> discrete_vertex(const Model<T>& m, int g, V3<T> pos, Q4<T> quat, int iv) {
    V3<T> vl = v3(T(0), T(0), T(0));
    if (m.geom_type[g] == GT_BOX) {
        const T* d = &m.geom_data[7 * g];
        vl = v3(((iv & 1) ? T(1) : T(-1)) * d[0] * T(0.5), ((iv & 2) ? T(1) : T(-1)) * d[1] * T(0.5),
                ((iv & 4) ? T(1) : T(-1)) * d[2] * T(0.5));
    } else if (m.geom_type[g] == GT_MESH) {
        int k = m.geom_vert_start[g] + iv;
        vl = v3(m.verts[3 * k], m.verts[3 * k + 1], m.verts[3 * k + 2]);
    }
    return transform_by_trans_quat(vl, pos, quat);
}

template <class T> bool is_discrete(const Model<T>& m, int g) {
    return m.geom_type[g] == GT_MESH || m.geom_type[g] == GT_BOX;
}

// gjk.py:1551-1657 func_search_valid_simplex_vertex
template <class T>
bool search_valid_vertex(const Model<T>& m, GJKState<T>& s, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb,
                         V3<T>& o1, V3<T>& o2, int& id1, int& id2, V3<T>& mink) {
    o1 = o2 = mink = v3(T(0), T(0), T(0));
    id1 = id2 = -1;
    if (is_discrete(m, ga) && is_discrete(m, gb)) {
        int na = m.geom_vert_end[ga] - m.geom_vert_start[ga];
        int nb = m.geom_vert_end[gb] - m.geom_vert_start[gb];
        int ncase = na * nb;
        for (int k = 0; k < ncase; k++) {
            int mm = (k + s.last_searched) % ncase;
            int i = mm / nb, j = mm % nb;
            id1 = m.geom_vert_start[ga] + i;
            id2 = m.geom_vert_start[gb] + j;
            o1 = discrete_vertex(m, ga, pa, qa, i);
            o2 = discrete_vertex(m, gb, pb, qb, j);
            mink = o1 - o2;
            if (simplex_valid(s, id1, id2, mink)) {
                s.last_searched = (mm + 1) % ncase;
                return true;
            }
        }
        return false;
    }
    if (s.snv == 3) {
        V3<T> dir = normalized(cross(s.s_mink[2] - s.s_mink[0], s.s_mink[1] - s.s_mink[0]));
        for (int i = 0; i < 2; i++) {
            V3<T> d = i == 0 ? dir : -dir;
            gjk_support_pair(m, ga, gb, d, pa, qa, pb, qb, o1, o2, id1, id2, mink);
            if (simplex_valid(s, id1, id2, mink)) return true;
        }
    }
    return false;
}

// gjk.py:1745-1857 func_safe_gjk_support: the direction is perturbed by EPS steps until the support is unique or the
// simplex stays valid. The perturbation signs are the reference's: -(1 - 2 (j & bit)) with bit = 1, 2, 4.
template <class T>
void safe_gjk_support(const Model<T>& m, GJKState<T>& s, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb,
                      V3<T> dir, V3<T>& o1, V3<T>& o2, int& id1, int& id2, V3<T>& mink) {
    const T EPS = m.EPS;
    o1 = o2 = v3(T(0), T(0), T(0));
    id1 = id2 = -1;
    mink = o1 - o2;
    for (int i = 0; i < 9; i++) {
        V3<T> nd = dir;
        if (i > 0) {
            int j = i - 1;
            nd[0] += -(T(1) - T(2) * T(j & 1)) * EPS;
            nd[1] += -(T(1) - T(2) * T(j & 2)) * EPS;
            nd[2] += -(T(1) - T(2) * T(j & 4)) * EPS;
        }
        nd = nd * (T(2) - dot(nd, dir));
        int nsup = count_support_geom(m, nd, ga, qa) * count_support_geom(m, -nd, gb, qb);
        if (i > 0 && nsup > 1) continue;
        gjk_support_pair(m, ga, gb, nd, pa, qa, pb, qb, o1, o2, id1, id2, mink);
        if (i == 0) {
            if (nsup > 1) continue;
            break;
        }
        if (i == 8) break;
        if (simplex_valid(s, id1, id2, mink)) break;
    }
}

// gjk.py:1711-1742 func_safe_gjk_triangle_info
template <class T> void safe_triangle_info(const GJKState<T>& s, int ta, int tb, int tc, int apex, V3<T>& n, T& sd) {
    V3<T> v1 = s.s_mink[ta], v2 = s.s_mink[tb], v3_ = s.s_mink[tc], ap = s.s_mink[apex];
    n = normalized(cross(v3_ - v1, v2 - v1));
    if (dot(n, ap - v1) > T(0)) n = -n;
    sd = dot(n, v1);
}

template <class T> void simplex_set(GJKState<T>& s, int i, V3<T> o1, V3<T> o2, int id1, int id2, V3<T> mink) {
    s.s_obj1[i] = o1;
    s.s_obj2[i] = o2;
    s.s_id1[i] = id1;
    s.s_id2[i] = id2;
    s.s_mink[i] = mink;
}

// gjk.py:1209-1425 func_safe_gjk
template <class T>
int safe_gjk(const Model<T>& m, GJKState<T>& s, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb) {
    bool init_ok = true;
    s.snv = 0;
    for (int i = 0; i < 4; i++) {
        V3<T> dir = v3(T(0), T(0), T(0));
        dir[2 - i / 2] = T(1) - T(2) * T(i % 2);
        V3<T> o1, o2, mink;
        int id1, id2;
        safe_gjk_support(m, s, ga, gb, pa, qa, pb, qb, dir, o1, o2, id1, id2, mink);
        if (!simplex_valid(s, id1, id2, mink)) {
            if (!search_valid_vertex(m, s, ga, gb, pa, qa, pb, qb, o1, o2, id1, id2, mink)) {
                init_ok = false;
                break;
            }
        }
        simplex_set(s, i, o1, o2, id1, id2, mink);
        s.snv++;
    }
    int flag = GJK_SEPARATED;
    if (init_ok) {
        int si[4] = {0, 1, 2, 3};
        V3<T> fn[4];
        T fd[4];
        for (int it = 0; it < GJK_MAX_ITER; it++) {
            for (int j = 0; j < 4; j++) {
                int s0 = si[2], s1 = si[1], s2 = si[3], ap = si[0];
                if (j == 1) s0 = si[0], s1 = si[2], s2 = si[3], ap = si[1];
                else if (j == 2) s0 = si[1], s1 = si[0], s2 = si[3], ap = si[2];
                else if (j == 3) s0 = si[0], s1 = si[1], s2 = si[2], ap = si[3];
                safe_triangle_info(s, s0, s1, s2, ap, fn[j], fd[j]);
            }
            int mi = 0;
            for (int j = 1; j < 4; j++)
                if (fd[j] < fd[mi]) mi = j;
            int msi = si[mi];
            V3<T> mn = fn[mi];
            T md = fd[mi];
            if (md >= T(0)) {
                flag = GJK_INTERSECT;
                break;
            }
            s.snv = 3;
            if (msi != 3) simplex_set(s, msi, s.s_obj1[3], s.s_obj2[3], s.s_id1[3], s.s_id2[3], s.s_mink[3]);
            V3<T> o1, o2, mink;
            int id1, id2;
            safe_gjk_support(m, s, ga, gb, pa, qa, pb, qb, mn, o1, o2, id1, id2, mink);
            if (simplex_duplicate(s, id1, id2)) {
                flag = GJK_SEPARATED;
                break;
            }
            if (simplex_degenerate(s, mink)) {
                flag = GJK_NUM_ERROR;
                break;
            }
            if (dot(mink, mn) < T(0)) {
                flag = GJK_SEPARATED;
                break;
            }
            simplex_set(s, 3, o1, o2, id1, id2, mink);
            s.snv = 4;
        }
    }
    if (flag == GJK_INTERSECT) {
        s.distance = T(0);
    } else {
        flag = GJK_SEPARATED;
        s.distance = GJKInfo<T>::float_max();
    }
    return flag;
}

// ------------------------------------------------------------------------------------------------------ EPA

// epa.py:1388-1432 func_plane_normal
template <class T> int plane_normal(V3<T> v1, V3<T> v2, V3<T> v3_, V3<T>& normal) {
    normal = v3(T(0), T(0), T(0));
    V3<T> d21 = v2 - v1, d31 = v3_ - v1, d32 = v3_ - v2;
    for (int i = 0; i < 3; i++) {
        V3<T> n = i == 0 ? cross(d32, d21) : (i == 1 ? cross(d21, d31) : cross(d31, d32));
        T nn = norm(n);
        if (nn == T(0)) return RC_FAIL;
        if (nn > GJKInfo<T>::float_min()) {
            normal = normalized(n);
            return RC_SUCCESS;
        }
    }
    return RC_FAIL;
}

// epa.py:407-431 func_epa_insert_vertex_to_polytope
template <class T> int epa_insert_vertex(GJKState<T>& s, V3<T> o1, V3<T> o2, int id1, int id2, V3<T> mink) {
    int n = s.nverts;
    s.p_obj1[n] = o1;
    s.p_obj2[n] = o2;
    s.p_id1[n] = id1;
    s.p_id2[n] = id2;
    s.p_mink[n] = mink;
    s.nverts++;
    return n;
}

// epa.py:1303-1385 func_safe_attach_face_to_polytope
template <class T> int safe_attach_face(GJKState<T>& s, int v1, int v2, int v3_, int a1, int a2, int a3) {
    int n = s.nfaces;
    s.f_verts[n][0] = v1, s.f_verts[n][1] = v2, s.f_verts[n][2] = v3_;
    s.f_adj[n][0] = a1, s.f_adj[n][1] = a2, s.f_adj[n][2] = a3;
    s.nfaces++;
    V3<T> normal;
    int flag = plane_normal(s.p_mink[v3_], s.p_mink[v2], s.p_mink[v1], normal);
    if (flag == RC_SUCCESS) {
        V3<T> fc = (s.p_mink[v1] + s.p_mink[v2] + s.p_mink[v3_]) / T(3);
        T max_orient = -dot(normal, fc);
        T max_abs = std::fabs(max_orient);
        for (int iv = 0; iv < s.nverts; iv++) {
            if (iv != v1 && iv != v2 && iv != v3_) {
                T o = dot(normal, s.p_mink[iv] - fc);
                if (std::fabs(o) > max_abs) {
                    max_abs = std::fabs(o);
                    max_orient = o;
                }
            }
        }
        if (max_orient > T(0)) normal = -normal;
        s.f_normal[n] = normal;
        T min_d2 = GJKInfo<T>::float_max();
        for (int k = 0; k < 3; k++) {
            int iv = k == 0 ? v1 : (k == 1 ? v2 : v3_);
            T dd = dot(normal, s.p_mink[iv]);
            T d2 = dd * dd;
            if (d2 < min_d2) min_d2 = d2;
        }
        s.f_dist2[n] = min_d2;
        s.f_map_idx[n] = -1;
    }
    return flag;
}

// epa.py:1250-1300 func_safe_epa_init: the tetrahedron's four faces
template <class T> void safe_epa_init(GJKState<T>& s) {
    int vi[4];
    for (int i = 0; i < 4; i++) vi[i] = epa_insert_vertex(s, s.s_obj1[i], s.s_obj2[i], s.s_id1[i], s.s_id2[i], s.s_mink[i]);
    static const int fv[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {3, 2, 1}};
    static const int fa[4][3] = {{1, 3, 2}, {2, 3, 0}, {0, 3, 1}, {2, 0, 1}};
    for (int i = 0; i < 4; i++) safe_attach_face(s, vi[fv[i][0]], vi[fv[i][1]], vi[fv[i][2]], fa[i][0], fa[i][1], fa[i][2]);
    for (int i = 0; i < 4; i++) {
        s.faces_map[i] = i;
        s.f_map_idx[i] = i;
    }
    s.nfaces_map = 4;
}

// epa.py:383-404 func_delete_face_from_polytope
template <class T> void delete_face(GJKState<T>& s, int f) {
    int mi = s.f_map_idx[f];
    if (mi >= 0) {
        int last = s.faces_map[s.nfaces_map - 1];
        s.faces_map[mi] = last;
        s.f_map_idx[last] = mi;
        s.nfaces_map--;
    }
    s.f_map_idx[f] = -2;
}

// epa.py:361-380 func_get_edge_idx
template <class T> int edge_idx(const GJKState<T>& s, int f, int v) {
    if (s.f_verts[f][0] == v) return 0;
    if (s.f_verts[f][1] == v) return 1;
    return 2;
}

// epa.py:273-340 func_epa_horizon (depth-first walk of the faces visible from w); returns RC_FAIL when a capacity of
// the reference's arrays would be exceeded
template <class T> int epa_horizon(GJKState<T>& s, int nearest) {
    V3<T> w = s.horizon_w;
    s.hs_face[0] = nearest;
    s.hs_edge[0] = 0;
    int top = 1;
    bool first = true;
    while (top > 0) {
        int f = s.hs_face[top - 1], e = s.hs_edge[top - 1];
        V3<T> v = s.p_mink[s.f_verts[f][0]];
        top--;
        bool deleted = s.f_map_idx[f] == -2;
        if (!first && deleted) continue;
        bool visible = dot(s.f_normal[f], w - v) > GJKInfo<T>::float_min();
        if (visible || first) {
            delete_face(s, f);
            for (int k = first ? 0 : 1; k < 3; k++) {
                int e2 = (e + k) % 3;
                int adj = s.f_adj[f][e2];
                if (s.f_map_idx[adj] != -2) {
                    int sv = s.f_verts[f][(e2 + 1) % 3];
                    if (top >= EPA_MAX_STACK) return RC_FAIL;
                    s.hs_face[top] = adj;
                    s.hs_edge[top] = edge_idx(s, adj, sv);
                    top++;
                }
            }
        } else {
            if (s.h_nedges >= EPA_MAX_HORIZON) return RC_FAIL;
            s.hd_edge[s.h_nedges] = e;
            s.hd_face[s.h_nedges] = f;
            s.h_nedges++;
        }
        first = false;
    }
    return RC_SUCCESS;
}

// gjk_utils.py:189-235 func_project_origin_to_plane
template <class T> int project_origin_to_plane(V3<T> v1, V3<T> v2, V3<T> v3_, V3<T>& point) {
    point = v3(T(0), T(0), T(0));
    V3<T> d21 = v2 - v1, d31 = v3_ - v1, d32 = v3_ - v2;
    for (int i = 0; i < 3; i++) {
        V3<T> n, v;
        if (i == 0) n = cross(d32, d21), v = v2;
        else if (i == 1) n = cross(d21, d31), v = v1;
        else n = cross(d31, d32), v = v3_;
        T nv = dot(n, v), nn = norm_sqr(n);
        if (nn == T(0)) return RC_FAIL;
        if (nn > GJKInfo<T>::float_min()) {
            point = n * (nv / nn);
            return RC_SUCCESS;
        }
        if (i == 2) {
            if (nn < GJKInfo<T>::float_min()) return RC_FAIL;
            point = n * (nv / nn);
            return RC_SUCCESS;
        }
    }
    return RC_SUCCESS;
}

// gjk_utils.py:48-107 func_triangle_affine_coords
template <class T> V3<T> triangle_affine_coords(V3<T> p, V3<T> t1, V3<T> t2, V3<T> t3) {
    T ms[3];
    for (int i = 0; i < 3; i++) {
        int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
        if (i == 1) std::swap(i1, i2);
        ms[i] = t2[i1] * t3[i2] - t2[i2] * t3[i1] - t1[i1] * t3[i2] + t1[i2] * t3[i1] + t1[i1] * t2[i2] -
                t1[i2] * t2[i1];
    }
    T m_max = T(0);
    int ix = 0, iy = 0;
    for (int i = 0; i < 3; i++) {
        T a = std::fabs(ms[i]), b = std::fabs(ms[(i + 1) % 3]), c = std::fabs(ms[(i + 2) % 3]);
        if (a >= b && a >= c) {
            m_max = ms[i];
            ix = (i + 1) % 3, iy = (i + 2) % 3;
            if (i == 1) std::swap(ix, iy);
            break;
        }
    }
    T cs[3];
    for (int i = 0; i < 3; i++) {
        V3<T> a = t2, b = t3;
        if (i == 1) a = t3, b = t1;
        else if (i == 2) a = t1, b = t2;
        cs[i] = p[ix] * a[iy] + p[iy] * b[ix] + a[ix] * b[iy] - p[ix] * b[iy] - p[iy] * a[ix] - b[ix] * a[iy];
    }
    return v3(cs[0] / m_max, cs[1] / m_max, cs[2] / m_max);
}

// epa.py:1183-1247 func_safe_epa_witness
template <class T> int safe_epa_witness(GJKState<T>& s, int f) {
    int i1 = s.f_verts[f][0], i2 = s.f_verts[f][1], i3 = s.f_verts[f][2];
    V3<T> v1 = s.p_mink[i1], v2 = s.p_mink[i2], v3_ = s.p_mink[i3];
    V3<T> proj;
    project_origin_to_plane(v1, v2, v3_, proj);
    V3<T> lam = triangle_affine_coords(proj, v1, v2, v3_);
    V3<T> rp = v1 * lam[0] + v2 * lam[1] + v3_ * lam[2];
    T err = norm(proj - rp);
    T e12 = norm_sqr(v1 - v2), e23 = norm_sqr(v2 - v3_), e31 = norm_sqr(v3_ - v1);
    T mx = std::fmax(std::fmax(e12, e23), std::fmax(e31, GJKInfo<T>::float_min_sq()));
    T rel = err * (T(1) / std::sqrt(mx));
    if (rel > GJKInfo<T>::max_rel_reproj() && err > GJKInfo<T>::max_abs_reproj()) return RC_FAIL;
    s.w1 = s.p_obj1[i1] * lam[0] + s.p_obj1[i2] * lam[1] + s.p_obj1[i3] * lam[2];
    s.w2 = s.p_obj2[i1] * lam[0] + s.p_obj2[i2] * lam[1] + s.p_obj2[i3] * lam[2];
    return RC_SUCCESS;
}

// epa.py:969-1180 func_safe_epa
template <class T>
int safe_epa(const Model<T>& m, GJKState<T>& s, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb) {
    T upper = GJKInfo<T>::float_max(), upper2 = GJKInfo<T>::float_max_sq(), lower = T(0);
    T tol = GJKInfo<T>::tolerance();
    const T EPS = m.EPS;
    int nearest = -1, prev = -1;
    bool discrete = is_discrete(m, ga) && is_discrete(m, gb);
    if (discrete) tol = EPS;
    for (int k = 0; k < EPA_MAX_ITER; k++) {
        prev = nearest;
        T lower2 = GJKInfo<T>::float_max_sq();
        for (int i = 0; i < s.nfaces_map; i++) {
            int f = s.faces_map[i];
            if (s.f_dist2[f] < lower2) {
                lower2 = s.f_dist2[f];
                nearest = f;
            }
        }
        if (lower2 > upper2 || nearest == -1) {
            nearest = prev;
            break;
        }
        lower = std::sqrt(lower2);
        V3<T> dir = s.f_normal[nearest];
        // epa.py:809-882 func_epa_support with dir_norm = 1
        V3<T> o1, o2, mk;
        int id1, id2;
        gjk_support_pair(m, ga, gb, dir, pa, qa, pb, qb, o1, o2, id1, id2, mk);
        int wi = epa_insert_vertex(s, o1, o2, id1, id2, mk);
        V3<T> w = s.p_mink[wi];
        T upk = dot(w, dir);
        if (upk < upper) {
            upper = upk;
            upper2 = upper * upper;
        }
        if (upper - lower < tol) break;
        if (discrete) {
            bool rep = false;
            for (int i = 0; i < s.nverts; i++) {
                if (i == wi) continue;
                if (s.p_id1[i] == s.p_id1[wi] && s.p_id2[i] == s.p_id2[wi]) {
                    rep = true;
                    break;
                }
            }
            if (rep) break;
        }
        s.horizon_w = w;
        if (epa_horizon(s, nearest) != RC_SUCCESS) {
            nearest = -1;
            break;
        }
        if (s.h_nedges < 3) {
            nearest = -1;
            break;
        }
        int nfaces = s.nfaces, nedges = s.h_nedges;
        if (nfaces + nedges >= EPA_MAX_FACES) break;
        int aflag = RC_SUCCESS;
        for (int i = 0; i < nedges; i++) {
            int f0 = nfaces + i;
            int f1 = nfaces + (i + 1) % nedges;
            int hf = s.hd_face[i], he = s.hd_edge[i];
            int hv1 = s.f_verts[hf][he];
            int hv2 = s.f_verts[hf][(he + 1) % 3];
            s.f_adj[hf][he] = f0;
            int adj0 = i > 0 ? f0 - 1 : nfaces + nedges - 1;
            aflag = safe_attach_face(s, wi, hv2, hv1, f1, hf, adj0);
            if (aflag != RC_SUCCESS) break;
            T d2 = s.f_dist2[s.nfaces - 1];
            if (d2 >= lower2 - EPS && d2 <= upper2 + EPS) {
                int nm = s.nfaces_map;
                s.faces_map[nm] = f0;
                s.f_map_idx[f0] = nm;
                s.nfaces_map++;
            }
        }
        if (aflag != RC_SUCCESS) {
            nearest = -1;
            break;
        }
        s.h_nedges = 0;
        if (s.nfaces_map == 0 || nearest == -1) {
            nearest = -1;
            break;
        }
        if (s.nverts >= EPA_MAX_VERTS) break;  // the next iteration would insert past the vertex array
    }
    if (nearest != -1) {
        T d2 = s.f_dist2[nearest];
        if (safe_epa_witness(s, nearest) == RC_SUCCESS) {
            s.n_witness = 1;
            s.distance = -std::sqrt(d2);
        } else {
            s.n_witness = 0;
            s.distance = T(0);
            nearest = -1;
        }
    } else {
        s.n_witness = 0;
        s.distance = T(0);
    }
    return nearest;
}

// gjk.py:170-446 func_gjk_contact (default path): returns whether the geoms collide, with the contact normal (from a
// to b), position (midpoint of the witness points) and penetration
template <class T>
bool gjk_contact(const Model<T>& m, GJKState<T>& s, int ga, int gb, V3<T> pa, Q4<T> qa, V3<T> pb, Q4<T> qb,
                 V3<T>& normal, V3<T>& pos, T& pen) {
    s.last_searched = 0;
    s.n_witness = 0;
    int flag = safe_gjk(m, s, ga, gb, pa, qa, pb, qb);
    if (flag == GJK_INTERSECT) {
        s.nverts = 0;
        s.nfaces = 0;
        s.nfaces_map = 0;
        s.h_nedges = 0;
        safe_epa_init(s);
        safe_epa(m, s, ga, gb, pa, qa, pb, qb);
    }
    bool is_col = s.distance < T(0);
    pen = is_col ? -s.distance : T(0);
    int n_contacts = 0;
    if (is_col) {
        for (int i = 0; i < s.n_witness; i++) {
            V3<T> cp = T(0.5) * (s.w1 + s.w2);
            V3<T> n = s.w2 - s.w1;
            T nl = norm(n);
            if (nl < GJKInfo<T>::float_min()) continue;
            if (n_contacts == 0) {
                pos = cp;
                normal = n / nl;
            }
            n_contacts++;
        }
    }
    if (n_contacts == 0) {
        is_col = false;
        pen = T(0);
    }
    return is_col;
}

}  // namespace orc
