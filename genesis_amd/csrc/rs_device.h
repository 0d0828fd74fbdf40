// Device-side views and FP32 spatial algebra for the gfx950 rigid-step kernels.
// The math restates genesis/utils/geom.py:19-425 in the reference's evaluation order (so that FP32 results stay within
// rounding of the reference); everything here is float, with IEEE division and square root.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rigidstep.h"
#include "../../include/rigidstep_arrays.h"
#include "rs_fields.h"

namespace rs {

// ------------------------------------------------------------------------------------------------- model view
// Model arrays are read-only while kernels run, so they live in the constant address space: reads at addresses that
// are the same on every lane become scalar loads.
#define RS_CONST __attribute__((address_space(4)))
typedef const float RS_CONST* cfptr;
typedef const int32_t RS_CONST* ciptr;

struct DModel {
    int L, J, D, Q, G, E, NE, P, NSUP, KMAX, CMAX, BROAD, NCPP;
    float dt, gx, gy, gz, tolerance, ls_tolerance, meaninertia, mc_perturbation, mc_tolerance, eps;
    int iterations, ls_iterations, enable_collision, enable_joint_limit, disable_constraint, enable_multi_contact;
    int box_box_detection, integrator, solver_type;
    float mpr_to_gjk_overlap_ratio;
    int64_t work_stride;  // floats of the work field per environment
    int off_gjk;          // offset of the per-stage kernels' GJK/EPA polytope in an environment's work slice
#define RS_DM_PTR(name, is_int, kind, width) const typename std::conditional<is_int, int32_t, float>::type RS_CONST* name;
    RS_MODEL_ARRAYS(RS_DM_PTR)
#undef RS_DM_PTR
};

// ------------------------------------------------------------------------------------------------- state view
struct DState {
    int B;       // number of environments (plane stride)
    int64_t LS;  // > 0: per-environment (local-kind) fields are env-major records of LS elements; 0: [row][env] planes
#define RS_DS_PTR(name, is_int, kind, width) typename std::conditional<is_int, int32_t, float>::type* name;
    RS_STATE_FIELDS(RS_DS_PTR)
#undef RS_DS_PTR
};

// Field identifiers, in directory order, and their element types and count kinds.
enum FieldId {
#define RS_FID(name, is_int, kind, width) FID_##name,
    RS_STATE_FIELDS(RS_FID)
#undef RS_FID
    FID_COUNT
};
template <int F>
struct FieldOf;
#define RS_FOF(name, is_int, kind, width)                                                          \
    template <>                                                                                     \
    struct FieldOf<FID_##name> {                                                                    \
        using T = typename std::conditional<is_int, int32_t, float>::type;                          \
        static constexpr int count_kind = kind;                                                     \
        __device__ __forceinline__ static T* get(const DState& s) { return s.name; }               \
    };
RS_STATE_FIELDS(RS_FOF)
#undef RS_FOF

// kinds whose fields the fused kernel keeps in an environment's private region
__host__ __device__ constexpr bool is_local_kind(int kind) {
    return kind == RS_N_ONE || kind == RS_N_LINKS || kind == RS_N_JOINTS || kind == RS_N_DOFS || kind == RS_N_QS ||
           kind == RS_N_GEOMS || kind == RS_N_ENTITIES || kind == RS_N_DOFS2 || kind == RS_N_SORT;
}

// One environment's private-region fields (fused kernel): field F starts at base + off[F] and has row stride 1. The
// offsets stay in kernel-argument (constant) memory and every access computes its address from them, so the view
// holds two uniform values instead of a pointer per field.
struct LState {
    static constexpr int B = 1;
    float* base;
    const RS_CONST int* off;
};

template <int F>
__device__ __forceinline__ typename FieldOf<F>::T* fptr(const DState& s) {
    return FieldOf<F>::get(s);
}
template <int F>
__device__ __forceinline__ typename FieldOf<F>::T* fptr(const LState& s) {
    static_assert(is_local_kind(FieldOf<F>::count_kind), "a local view reaches private-region fields only");
    return (typename FieldOf<F>::T*)(s.base + s.off[F]);
}
// pointer to field `field` of state s (a DState: the field's first element; an LState: the private-region copy)
#define FP(s, field) fptr<FID_##field>(s)

// Environment b's elements of a field: element r at p[r * rs]. In a DState a local-kind field is either a plane
// ([row][env], rs = B) or, with env-major records (LS > 0), a slice of environment b's record (rs = 1); every other
// field is a plane. In an LState every field is a private-region slice.
template <class T>
struct FView {
    T* p;
    int rs;
    __device__ __forceinline__ T& operator[](int64_t r) const { return p[r * rs]; }
};
template <int F>
__device__ __forceinline__ FView<typename FieldOf<F>::T> field_view(const DState& s, int b) {
    typename FieldOf<F>::T* p = fptr<F>(s);
    if constexpr (is_local_kind(FieldOf<F>::count_kind)) {
        if (s.LS > 0) return {p + (size_t)b * s.LS, 1};
    }
    return {p + b, s.B};
}
template <int F>
__device__ __forceinline__ FView<typename FieldOf<F>::T> field_view(const LState& s, int) {
    return {fptr<F>(s), 1};
}
#define FV(s, field, b) field_view<FID_##field>(s, b)

struct f3 {
    float x, y, z;
    __device__ __forceinline__ float& operator[](int i) { return (&x)[i]; }
    __device__ __forceinline__ float operator[](int i) const { return (&x)[i]; }
};
struct f4q {
    float w, x, y, z;
};
struct m33 {
    float a[3][3];
};

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 zero3() { return f3{0.f, 0.f, 0.f}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float nsq3(f3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__device__ __forceinline__ float norm3(f3 a) { return sqrtf(nsq3(a)); }
__device__ __forceinline__ f3 normalize3(f3 a) {
    float inv = 1.0f / norm3(a);
    return inv * a;
}
__device__ __forceinline__ float nsq4(f4q q) { return q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z; }
__device__ __forceinline__ f4q normalize4(f4q q) {
    float inv = 1.0f / sqrtf(nsq4(q));
    return {inv * q.w, inv * q.x, inv * q.y, inv * q.z};
}
__device__ __forceinline__ f4q qinv(f4q q) { return {q.w, -q.x, -q.y, -q.z}; }
__device__ __forceinline__ f4q qident() { return {1.f, 0.f, 0.f, 0.f}; }

// u * v (Hamilton product written as in geom.py qd_quat_mul)
__device__ __forceinline__ f4q qmul(f4q u, f4q v) {
    return {u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z, u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y,
            u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x, u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w};
}
// geom.py:224-232 qd_quat_mul_axis: q times the pure quaternion (0, a)
__device__ __forceinline__ f4q qmul_axis(f4q q, f3 a) {
    return {-q.x * a.x - q.y * a.y - q.z * a.z, q.w * a.x + q.y * a.z - q.z * a.y, q.w * a.y + q.z * a.x - q.x * a.z,
            q.w * a.z + q.x * a.y - q.y * a.x};
}
// transform quat v by quat u: normalize(u * v)
__device__ __forceinline__ f4q qtrans(f4q v, f4q u) { return normalize4(qmul(u, v)); }

__device__ __forceinline__ f3 qrot(f3 v, f4q q) {
    float qxx = q.x * q.x, qxy = q.x * q.y, qxz = q.x * q.z, qwx = q.x * q.w;
    float qyy = q.y * q.y, qyz = q.y * q.z, qwy = q.y * q.w;
    float qzz = q.z * q.z, qwz = q.z * q.w, qww = q.w * q.w;
    float r0 = v.x * (qxx + qww - qyy - qzz) + v.y * (2.f * qxy - 2.f * qwz) + v.z * (2.f * qxz + 2.f * qwy);
    float r1 = v.x * (2.f * qwz + 2.f * qxy) + v.y * (qww - qxx + qyy - qzz) + v.z * (-2.f * qwx + 2.f * qyz);
    float r2 = v.x * (-2.f * qwy + 2.f * qxz) + v.y * (2.f * qwx + 2.f * qyz) + v.z * (qww - qxx - qyy + qzz);
    float d = qww + qxx + qyy + qzz;
    return {r0 / d, r1 / d, r2 / d};
}
__device__ __forceinline__ f3 qrot_inv(f3 v, f4q q) { return qrot(v, qinv(q)); }
__device__ __forceinline__ f3 tr_point(f3 p, f3 t, f4q q) { return qrot(p, q) + t; }
__device__ __forceinline__ void tr_pose(f3 pos, f4q quat, f3 tt, f4q tq, f3& np, f4q& nq) {
    np = tt + qrot(pos, tq);
    nq = qtrans(quat, tq);
}

__device__ __forceinline__ f4q rotvec_quat(f3 rv, float eps) {
    float th2 = nsq3(rv);
    f4q q = qident();
    if (th2 > eps * eps) {
        float th = sqrtf(th2);
        float half = 0.5f * th;
        float c = cosf(half), s = sinf(half);
        float k = s / th;
        q = {c, k * rv.x, k * rv.y, k * rv.z};
        float f = 0.5f * (3.f - nsq4(q));
        q = {q.w * f, q.x * f, q.y * f, q.z * f};
    }
    return q;
}

__device__ __forceinline__ m33 quat_R(f4q q, float eps) {
    m33 R{{{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}}};
    float d = nsq4(q);
    if (d > eps) {
        float s = 2.f / d;
        float xs = q.x * s, ys = q.y * s, zs = q.z * s;
        float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
        float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
        float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
        R = m33{{{1.f - (yy + zz), xy - wz, xz + wy}, {xy + wz, 1.f - (xx + zz), yz - wx},
                 {xz - wy, yz + wx, 1.f - (xx + yy)}}};
    }
    return R;
}

__device__ __forceinline__ f3 quat_xyz(f4q q, float eps) {
    float roll = 0.f, pitch = 0.f, yaw = 0.f;
    float n = nsq4(q);
    if (n > eps) {
        float s = 2.f / n;
        float xs = q.x * s, ys = q.y * s, zs = q.z * s;
        float wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
        float xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
        float yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
        float sy = wz - xy, cy = 1.f - (yy + zz);
        float cp = sqrtf(cy * cy + sy * sy);
        pitch = atan2f(xz + wy, cp);
        if (cp > eps) {
            roll = atan2f(wx - yz, 1.f - (xx + yy));
            yaw = atan2f(sy, cy);
        } else {
            yaw = atan2f(wz + xy, 1.f - (xx + zz));
        }
    }
    return {roll, pitch, yaw};
}

__device__ __forceinline__ f3 mv33(const m33& A, f3 v) {
    return {A.a[0][0] * v.x + A.a[0][1] * v.y + A.a[0][2] * v.z, A.a[1][0] * v.x + A.a[1][1] * v.y + A.a[1][2] * v.z,
            A.a[2][0] * v.x + A.a[2][1] * v.y + A.a[2][2] * v.z};
}
__device__ __forceinline__ m33 mm33(const m33& A, const m33& B) {
    m33 C;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C.a[i][j] = A.a[i][0] * B.a[0][j] + A.a[i][1] * B.a[1][j] + A.a[i][2] * B.a[2][j];
    return C;
}
__device__ __forceinline__ m33 tr33(const m33& A) {
    m33 C;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) C.a[i][j] = A.a[j][i];
    return C;
}

// spatial inertia of a body (i_inertial, mass) offset by trans and rotated by quat (geom.py:321-338)
__device__ __forceinline__ void inertia_offset(const m33& I, float mass, f3 t, f4q q, float eps, m33& oi, f3& op) {
    float xx = t.x * t.x, xy = t.x * t.y, xz = t.x * t.z, yy = t.y * t.y, yz = t.y * t.z, zz = t.z * t.z;
    m33 h{{{yy + zz, -xy, -xz}, {-xy, xx + zz, -yz}, {-xz, -yz, xx + yy}}};
    m33 R = quat_R(q, eps);
    m33 RI = mm33(mm33(R, I), tr33(R));
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) oi.a[i][j] = RI.a[i][j] + h.a[i][j] * mass;
    op = t * mass;
}
__device__ __forceinline__ void inertial_mul(f3 pos, const m33& I, float mass, f3 vel, f3 ang, f3& oa, f3& ov) {
    oa = mv33(I, ang) + cross3(pos, vel);
    ov = mass * vel - cross3(pos, ang);
}
__device__ __forceinline__ void motion_x_force(f3 ma, f3 mv, f3 fa, f3 fv, f3& oa, f3& ov) {
    ov = cross3(ma, fv);
    oa = cross3(ma, fa) + cross3(mv, fv);
}
__device__ __forceinline__ void motion_x_motion(f3 sa, f3 sv, f3 ma, f3 mv, f3& oa, f3& ov) {
    ov = cross3(sa, mv) + cross3(sv, ma);
    oa = cross3(sa, ma);
}
__device__ __forceinline__ void orthogonals(f3 a, f3& o1, f3& o2) {
    f3 b = zero3();
    if (fabsf(a.y) < 0.5f) {
        b.x = -a.x * a.y;
        b.y = 1.f - a.y * a.y;
        b.z = -a.z * a.y;
    } else {
        b.x = -a.x * a.z;
        b.y = -a.y * a.z;
        b.z = 1.f - a.z * a.z;
    }
    b = normalize3(b);
    o1 = cross3(b, a);
    o2 = b;
}
// impedance of a constraint at signed depth neg_pen (depends on the solver parameters and the depth only)
template <class P>
__device__ __forceinline__ float impedance(P p, float neg_pen) {
    float dmin = p[2], dmax = p[3], width = p[4], mid = p[5], power = p[6];
    float x = fabsf(neg_pen) / width;
    float ia = (1.f / powf(mid, power - 1.f)) * powf(x, power);
    float ib = 1.f - (1.f / powf(1.f - mid, power - 1.f)) * powf(1.f - x, power);
    float iy = x < mid ? ia : ib;
    float imp = dmin + iy * (dmax - dmin);
    imp = fminf(fmaxf(imp, dmin), dmax);
    return x > 1.f ? dmax : imp;
}
// reference acceleration of a constraint of impedance imp at velocity vel and position pos
template <class P>
__device__ __forceinline__ float reference_acc(P p, float imp, float vel, float pos) {
    float tc = p[0], dr = p[1], dmax = p[3];
    float b = 2.f / (dmax * tc);
    float k = 1.f / (dmax * dmax * tc * tc * dr * dr);
    return -b * vel - k * imp * pos;
}
template <class P>
__device__ __forceinline__ void imp_aref(P p, float neg_pen, float vel, float pos, float& imp, float& aref) {
    imp = impedance(p, neg_pen);
    aref = reference_acc(p, imp, vel, pos);
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

// ------------------------------------------------------------------------------------- model accessors
__device__ __forceinline__ f3 mvec3(cfptr a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
__device__ __forceinline__ f4q mquat(cfptr a, int i) { return {a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}; }
__device__ __forceinline__ m33 mmat3(cfptr a, int i) {
    m33 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.a[k / 3][k % 3] = a[9 * i + k];
    return r;
}

// --------------------------------------------------------------------------------- state plane accessors
// element (row) of a field for environment b: field[row * B + b]
#define SF(field, row) FV(st, field, b)[(row)]

__device__ __forceinline__ f3 ld3(const float* f, int i, int B, int b) {
    return {f[(size_t)(3 * i) * B + b], f[(size_t)(3 * i + 1) * B + b], f[(size_t)(3 * i + 2) * B + b]};
}
__device__ __forceinline__ void st3(float* f, int i, int B, int b, f3 v) {
    f[(size_t)(3 * i) * B + b] = v.x;
    f[(size_t)(3 * i + 1) * B + b] = v.y;
    f[(size_t)(3 * i + 2) * B + b] = v.z;
}
__device__ __forceinline__ f4q ld4(const float* f, int i, int B, int b) {
    return {f[(size_t)(4 * i) * B + b], f[(size_t)(4 * i + 1) * B + b], f[(size_t)(4 * i + 2) * B + b],
            f[(size_t)(4 * i + 3) * B + b]};
}
__device__ __forceinline__ void st4(float* f, int i, int B, int b, f4q q) {
    f[(size_t)(4 * i) * B + b] = q.w;
    f[(size_t)(4 * i + 1) * B + b] = q.x;
    f[(size_t)(4 * i + 2) * B + b] = q.y;
    f[(size_t)(4 * i + 3) * B + b] = q.z;
}
__device__ __forceinline__ m33 ld9(const float* f, int i, int B, int b) {
    m33 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.a[k / 3][k % 3] = f[(size_t)(9 * i + k) * B + b];
    return r;
}
__device__ __forceinline__ void st9(float* f, int i, int B, int b, const m33& m) {
#pragma unroll
    for (int k = 0; k < 9; k++) f[(size_t)(9 * i + k) * B + b] = m.a[k / 3][k % 3];
}
template <class T>
__device__ __forceinline__ f3 ld3(FView<T> f, int i) {
    return {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
}
template <class T>
__device__ __forceinline__ void st3(FView<T> f, int i, f3 v) {
    f[3 * i] = v.x;
    f[3 * i + 1] = v.y;
    f[3 * i + 2] = v.z;
}
template <class T>
__device__ __forceinline__ f4q ld4(FView<T> f, int i) {
    return {f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]};
}
template <class T>
__device__ __forceinline__ void st4(FView<T> f, int i, f4q q) {
    f[4 * i] = q.w;
    f[4 * i + 1] = q.x;
    f[4 * i + 2] = q.y;
    f[4 * i + 3] = q.z;
}
template <class T>
__device__ __forceinline__ m33 ld9(FView<T> f, int i) {
    m33 r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.a[k / 3][k % 3] = f[9 * i + k];
    return r;
}
template <class T>
__device__ __forceinline__ void st9(FView<T> f, int i, const m33& m) {
#pragma unroll
    for (int k = 0; k < 9; k++) f[9 * i + k] = m.a[k / 3][k % 3];
}
#define LD3(field, i) ld3(FV(st, field, b), (i))
#define ST3(field, i, v) st3(FV(st, field, b), (i), (v))
#define LD4(field, i) ld4(FV(st, field, b), (i))
#define ST4(field, i, v) st4(FV(st, field, b), (i), (v))
#define LD9(field, i) ld9(FV(st, field, b), (i))
#define ST9(field, i, v) st9(FV(st, field, b), (i), (v))

constexpr int CTRL_POSITION = 0, CTRL_VELOCITY = 1, CTRL_FORCE = 2;
constexpr int JT_FIXED = 0, JT_REVOLUTE = 1, JT_PRISMATIC = 2, JT_SPHERICAL = 3, JT_FREE = 4;
constexpr int GT_PLANE = 0, GT_SPHERE = 1, GT_ELLIPSOID = 2, GT_CYLINDER = 3, GT_CAPSULE = 4, GT_BOX = 5,
              GT_MESH = 6, GT_TERRAIN = 7;
constexpr int EQ_CONNECT = 0, EQ_WELD = 1, EQ_JOINT = 2;

}  // namespace rs
