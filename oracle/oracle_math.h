// ORACLE (test infrastructure only). Scalar restatement of the reference's quaternion/spatial algebra
// (genesis/utils/geom.py:19-425), templated on the floating type so the same source gives the FP32 per-stage
// reference and the FP64 drift oracle. Evaluation order follows the reference text where it changes FP32 rounding.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

template <class T> struct V3 {
    T x, y, z;
    T& operator[](int i) { return (&x)[i]; }
    const T& operator[](int i) const { return (&x)[i]; }
};
template <class T> struct Q4 {
    T w, x, y, z;
    T& operator[](int i) { return (&w)[i]; }
    const T& operator[](int i) const { return (&w)[i]; }
};
template <class T> struct M3 {
    T m[3][3];
};

template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T norm_sqr(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class T> inline T norm(V3<T> a) { return std::sqrt(norm_sqr(a)); }
// Matrix.normalized(): multiply by the reciprocal of the norm.
template <class T> inline V3<T> normalized(V3<T> a) {
    T inv = T(1) / norm(a);
    return inv * a;
}
template <class T> inline T norm_sqr(Q4<T> q) { return q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z; }
template <class T> inline Q4<T> normalized(Q4<T> q) {
    T inv = T(1) / std::sqrt(norm_sqr(q));
    return {inv * q.w, inv * q.x, inv * q.y, inv * q.z};
}
template <class T> inline Q4<T> identity_quat() { return {T(1), T(0), T(0), T(0)}; }
template <class T> inline Q4<T> inv_quat(Q4<T> q) { return {q.w, -q.x, -q.y, -q.z}; }

// geom.py:223-230 qd_quat_mul (outer-product form)
template <class T> inline Q4<T> quat_mul(Q4<T> u, Q4<T> v) {
    T w = u.w * v.w - u.x * v.x - u.y * v.y - u.z * v.z;
    T x = u.w * v.x + u.x * v.w + u.y * v.z - u.z * v.y;
    T y = u.w * v.y - u.x * v.z + u.y * v.w + u.z * v.x;
    T z = u.w * v.z + u.x * v.y - u.y * v.x + u.z * v.w;
    return {w, x, y, z};
}
// geom.py:246-252: quat_mul(u, v) normalised ("transform v by u")
template <class T> inline Q4<T> transform_quat_by_quat(Q4<T> v, Q4<T> u) { return normalized(quat_mul(u, v)); }
// geom.py:224-232 qd_quat_mul_axis: q times the pure quaternion (0, a)
template <class T> inline Q4<T> quat_mul_axis(Q4<T> q, V3<T> a) {
    return {-q.x * a.x - q.y * a.y - q.z * a.z, q.w * a.x + q.y * a.z - q.z * a.y, q.w * a.y + q.z * a.x - q.x * a.z,
            q.w * a.z + q.x * a.y - q.y * a.x};
}

// geom.py:256-270 (divides by |q|^2)
template <class T> inline V3<T> transform_by_quat(V3<T> v, Q4<T> q) {
    T qxx = q.x * q.x, qxy = q.x * q.y, qxz = q.x * q.z, qwx = q.x * q.w;
    T qyy = q.y * q.y, qyz = q.y * q.z, qwy = q.y * q.w;
    T qzz = q.z * q.z, qwz = q.z * q.w;
    T qww = q.w * q.w;
    T r0 = v.x * (qxx + qww - qyy - qzz) + v.y * (T(2) * qxy - T(2) * qwz) + v.z * (T(2) * qxz + T(2) * qwy);
    T r1 = v.x * (T(2) * qwz + T(2) * qxy) + v.y * (qww - qxx + qyy - qzz) + v.z * (T(-2) * qwx + T(2) * qyz);
    T r2 = v.x * (T(-2) * qwy + T(2) * qxz) + v.y * (T(2) * qwx + T(2) * qyz) + v.z * (qww - qxx - qyy + qzz);
    T d = qww + qxx + qyy + qzz;
    return {r0 / d, r1 / d, r2 / d};
}
template <class T> inline V3<T> inv_transform_by_quat(V3<T> v, Q4<T> q) { return transform_by_quat(v, inv_quat(q)); }
template <class T> inline V3<T> transform_by_trans_quat(V3<T> p, V3<T> t, Q4<T> q) {
    return transform_by_quat(p, q) + t;
}
template <class T> inline V3<T> inv_transform_by_trans_quat(V3<T> p, V3<T> t, Q4<T> q) {
    return transform_by_quat(p - t, inv_quat(q));
}
// geom.py:306-309
template <class T>
inline void transform_pos_quat_by_trans_quat(V3<T> pos, Q4<T> quat, V3<T> tt, Q4<T> tq, V3<T>& npos, Q4<T>& nquat) {
    npos = tt + transform_by_quat(pos, tq);
    nquat = transform_quat_by_quat(quat, tq);
}

// geom.py:111-133 (first-order renormalisation)
template <class T> inline Q4<T> rotvec_to_quat(V3<T> rv, T eps) {
    T thetasq = norm_sqr(rv);
    Q4<T> q{T(1), T(0), T(0), T(0)};
    if (thetasq > eps * eps) {
        T theta = std::sqrt(thetasq);
        T half = T(0.5) * theta;
        T c = std::cos(half), s = std::sin(half);
        T k = s / theta;
        q = {c, k * rv.x, k * rv.y, k * rv.z};
        T f = T(0.5) * (T(3) - norm_sqr(q));
        q = {q.w * f, q.x * f, q.y * f, q.z * f};
    }
    return q;
}

// geom.py:137-162
template <class T> inline M3<T> quat_to_R(Q4<T> q, T eps) {
    M3<T> R{{{T(1), T(0), T(0)}, {T(0), T(1), T(0)}, {T(0), T(0), T(1)}}};
    T d = norm_sqr(q);
    if (d > eps) {
        T s = T(2) / d;
        T xs = q.x * s, ys = q.y * s, zs = q.z * s;
        T wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
        T xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
        T yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
        R = M3<T>{{{T(1) - (yy + zz), xy - wz, xz + wy}, {xy + wz, T(1) - (xx + zz), yz - wx},
                   {xz - wy, yz + wx, T(1) - (xx + yy)}}};
    }
    return R;
}

// geom.py:166-191
template <class T> inline V3<T> quat_to_xyz(Q4<T> q, T eps) {
    T roll = 0, pitch = 0, yaw = 0;
    T n = norm_sqr(q);
    if (n > eps) {
        T s = T(2) / n;
        T xs = q.x * s, ys = q.y * s, zs = q.z * s;
        T wx = q.w * xs, wy = q.w * ys, wz = q.w * zs;
        T xx = q.x * xs, xy = q.x * ys, xz = q.x * zs;
        T yy = q.y * ys, yz = q.y * zs, zz = q.z * zs;
        T sinycosp = wz - xy;
        T cosycosp = T(1) - (yy + zz);
        T cosp = std::sqrt(cosycosp * cosycosp + sinycosp * sinycosp);
        pitch = std::atan2(xz + wy, cosp);
        if (cosp > eps) {
            roll = std::atan2(wx - yz, T(1) - (xx + yy));
            yaw = std::atan2(sinycosp, cosycosp);
        } else {
            yaw = std::atan2(wz + xy, T(1) - (xx + zz));
        }
    }
    return {roll, pitch, yaw};
}

template <class T> inline V3<T> matvec(const M3<T>& A, V3<T> v) {
    return {A.m[0][0] * v.x + A.m[0][1] * v.y + A.m[0][2] * v.z, A.m[1][0] * v.x + A.m[1][1] * v.y + A.m[1][2] * v.z,
            A.m[2][0] * v.x + A.m[2][1] * v.y + A.m[2][2] * v.z};
}
template <class T> inline M3<T> matmul(const M3<T>& A, const M3<T>& B) {
    M3<T> C;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
    return C;
}
template <class T> inline M3<T> transpose(const M3<T>& A) {
    M3<T> C;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C.m[i][j] = A.m[j][i];
    return C;
}

// geom.py:321-338: cinr = (R I R^T + hhT m, m*trans, quat, m)
template <class T>
inline void transform_inertia_by_trans_quat(const M3<T>& I, T mass, V3<T> trans, Q4<T> quat, T eps, M3<T>& out_i,
                                            V3<T>& out_pos) {
    T x = trans.x, y = trans.y, z = trans.z;
    T xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    M3<T> hhT{{{yy + zz, -xy, -xz}, {-xy, xx + zz, -yz}, {-xz, -yz, xx + yy}}};
    M3<T> R = quat_to_R(quat, eps);
    M3<T> RI = matmul(matmul(R, I), transpose(R));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out_i.m[i][j] = RI.m[i][j] + hhT.m[i][j] * mass;
    out_pos = trans * mass;
}

// geom.py:366-369
template <class T>
inline void inertial_mul(V3<T> pos, const M3<T>& i, T mass, V3<T> vel, V3<T> ang, V3<T>& o_ang, V3<T>& o_vel) {
    o_ang = matvec(i, ang) + cross(pos, vel);
    o_vel = mass * vel - cross(pos, ang);
}
// geom.py:373-377
template <class T>
inline void motion_cross_force(V3<T> m_ang, V3<T> m_vel, V3<T> f_ang, V3<T> f_vel, V3<T>& o_ang, V3<T>& o_vel) {
    o_vel = cross(m_ang, f_vel);
    o_ang = cross(m_ang, f_ang) + cross(m_vel, f_vel);
}
// geom.py:380-384
template <class T>
inline void motion_cross_motion(V3<T> s_ang, V3<T> s_vel, V3<T> m_ang, V3<T> m_vel, V3<T>& o_ang, V3<T>& o_vel) {
    o_vel = cross(s_ang, m_vel) + cross(s_vel, m_ang);
    o_ang = cross(s_ang, m_ang);
}

// geom.py:387-404
template <class T> inline void orthogonals(V3<T> a, V3<T>& o1, V3<T>& o2) {
    V3<T> b{T(0), T(0), T(0)};
    if (std::fabs(a.y) < T(0.5)) {
        b.x = -a.x * a.y;
        b.y = T(1) - a.y * a.y;
        b.z = -a.z * a.y;
    } else {
        b.x = -a.x * a.z;
        b.y = -a.y * a.z;
        b.z = T(1) - a.z * a.z;
    }
    b = normalized(b);
    o1 = cross(b, a);
    o2 = b;
}

// geom.py:407-425 (MuJoCo solref/solimp impedance model)
template <class T> inline void imp_aref(const T* p, T neg_pen, T vel, T pos, T& imp, T& aref) {
    T timeconst = p[0], dampratio = p[1], dmin = p[2], dmax = p[3], width = p[4], mid = p[5], power = p[6];
    T imp_x = std::fabs(neg_pen) / width;
    T imp_a = (T(1) / std::pow(mid, power - T(1))) * std::pow(imp_x, power);
    T imp_b = T(1) - (T(1) / std::pow(T(1) - mid, power - T(1))) * std::pow(T(1) - imp_x, power);
    T imp_y = imp_x < mid ? imp_a : imp_b;
    imp = dmin + imp_y * (dmax - dmin);
    imp = std::fmin(std::fmax(imp, dmin), dmax);
    imp = imp_x > T(1) ? dmax : imp;
    T b = T(2) / (dmax * timeconst);
    T k = T(1) / (dmax * dmax * timeconst * timeconst * dampratio * dampratio);
    aref = -b * vel - k * imp * pos;
}

template <class T> inline T clampv(T x, T lo, T hi) { return std::fmin(std::fmax(x, lo), hi); }

}  // namespace orc
