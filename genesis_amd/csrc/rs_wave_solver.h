// Constraint rows and Newton solve of one environment on one wavefront (fused substep).
// Same algorithm and arithmetic as rs_constraint.h (constraint/solver.py: rows :580-1286, solve :2936-3261, line
// search :2034-2597): row-wise work is split across lanes, and every reduction over rows or dofs is summed in the
// reference's sequential order (by one lane per output, or uniformly by all lanes), so results match the per-stage
// kernels. Rows live in the env-major work field (Rows); the small matrices (mass matrix, Hessian factor) and dof
// vectors live in the environment's private region (loc).
#pragma once
#include "rs_constraint.h"
#include "rs_wave.h"

namespace rs {

// Row store of the fused kernel: env-major arrays in the work field.
struct WorkRows {
    float* J;   // [c * D + d]
    float* f;   // [k * cap + c], k < RW_COUNT
    float* tp;  // [k * cap + c], k < kTmpArrays
    int cap, D;
    __device__ __forceinline__ float& jac(int c, int d) const { return J[(size_t)c * D + d]; }
    __device__ __forceinline__ float& diag(int c) const { return f[(size_t)RW_DIAG * cap + c]; }
    __device__ __forceinline__ float& aref(int c) const { return f[(size_t)RW_AREF * cap + c]; }
    __device__ __forceinline__ float& efc_D(int c) const { return f[(size_t)RW_EFC_D * cap + c]; }
    __device__ __forceinline__ float& fl(int c) const { return f[(size_t)RW_FL * cap + c]; }
    __device__ __forceinline__ float& Jaref(int c) const { return f[(size_t)RW_JAREF * cap + c]; }
    __device__ __forceinline__ float& jv(int c) const { return f[(size_t)RW_JV * cap + c]; }
    __device__ __forceinline__ float& force(int c) const { return f[(size_t)RW_FORCE * cap + c]; }
    __device__ __forceinline__ int& active(int c) const { return ((int*)f)[(size_t)RW_ACTIVE * cap + c]; }
    __device__ __forceinline__ int& prev(int c) const { return ((int*)f)[(size_t)RW_PREV * cap + c]; }
    __device__ __forceinline__ float& tmp(int k, int c) const { return tp[(size_t)k * cap + c]; }
};

#define LF(field, row) FP(loc, field)[(row)]

// a lane's column of the row tile as the staging buffer of the contact jacobian rows it builds
struct TileJac {
    float* col;
    __device__ __forceinline__ float& at(const WorkRows&, int, int k) const { return col[k * kTileStride]; }
    __device__ __forceinline__ void flush(const WorkRows& rows, int c, int D) const {
        for (int k = 0; k < D; k++) rows.jac(c, k) = col[k * kTileStride];
    }
};

// solver.py:580-1286: equality, friction-loss and limit rows uniformly (few rows), contact rows one lane per row (with a
// row tile) or per contact
__device__ inline void wave_constraints(const DModel& m, const LState& loc, const DState& st, int b, const WorkRows& R,
                                        int lane, float* tile) {
    int nc = equality_rows(m, loc, 0, R);
    const int neq = nc;
    int nfl = frictionloss_rows(m, loc, 0, R, nc);
    nc += nfl;
    if (m.enable_collision) {
        const int ncont = LF(n_contacts, 0);
        if (tile) {  // one lane per row: lane 4 q + i builds row i of the q-th contact of the pass
            const TileJac jt{tile + lane};
            // rows D .. D + 2 of the tile are free until the solve: the walk's tree arrays go there when they fit
            const int D = m.D, L = m.L;
            if (3 * L <= 3 * kTileStride && tile_rows(D) >= D + 3) {
                int* tree = (int*)(tile + D * kTileStride);
                for (int l = lane; l < L; l += kWave) {
                    tree[l] = m.link_parent[l];
                    tree[L + l] = m.link_dof_start[l];
                    tree[2 * L + l] = m.link_dof_end[l];
                }
                wave_sync();
                const ChainLds ch{tree, L};
                for (int r = lane; r < 4 * ncont; r += kWave)
                    contact_row(m, loc, 0, st, b, R, r >> 2, r & 3, nc, jt, ch);
            } else {
                for (int r = lane; r < 4 * ncont; r += kWave) contact_row(m, loc, 0, st, b, R, r >> 2, r & 3, nc, jt);
            }
        } else {
            for (int ic = lane; ic < ncont; ic += kWave) contact_rows(m, loc, 0, st, b, R, ic, nc);
        }
        nc += ncont * 4;
    }
    if (m.enable_joint_limit) nc = limit_rows(m, loc, 0, R, nc);
    LF(n_con, 0) = nc;
    LF(n_con_eq, 0) = neq;
    LF(n_con_fl, 0) = nfl;
    wave_sync();
}

typedef float mfma_v4f __attribute__((ext_vector_type(4)));
constexpr int kMfmaMaxDofs = 128;  // 8 x 8 tiles of 16
// models above the row-tile classes run their dense row products on the matrix cores
__device__ __forceinline__ bool mfma_model(int D) { return D > kTileWideDofs && D <= kMfmaMaxDofs; }

// Sum over the 64 lanes with DPP (row_shr 1, 2, 4, 8 within each row of 16, then row_bcast:15 and row_bcast:31):
// a source lane outside the row, or a row left out by the row mask, contributes 0; lane 63 ends with the sum.
template <int kCtrl, int kRowMask = 0xf>
__device__ __forceinline__ float dpp_zero_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, kRowMask, 0xf, false));
}
__device__ __forceinline__ float wave_sum_f(float x) {
    x = x + dpp_zero_f<0x111>(x);
    x = x + dpp_zero_f<0x112>(x);
    x = x + dpp_zero_f<0x114>(x);
    x = x + dpp_zero_f<0x118>(x);
    x = x + dpp_zero_f<0x142, 0xa>(x);
    x = x + dpp_zero_f<0x143, 0xc>(x);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), kWave - 1));
}

// sum over d < D of f(d) with lane l taking d = l, l + 64, ..., reduced across the wavefront (for the matrix-core
// models, whose private region is in global memory: a loop over d on every lane would wait on one load per d)
template <class F>
__device__ __forceinline__ float wave_dsum(int D, int lane, F f) {
    float p = 0.f;
    for (int d = lane; d < D; d += kWave) p = p + f(d);
    return wave_sum_f(p);
}

// out[d] = sum over rows c < nc of J[c][d] w[c] (d < D <= 128) on the matrix cores: for each 16-column tile, A holds
// the row weights (A[i][k] = w[c0 + k] for every i) and B the rows (B[k][j] = J[c0 + k][16 T + j]), so every row of
// the tile's result is the sum; 16 rows per trip, their operands loaded before the products.
template <class W>
__device__ inline void gemv_t_mfma(const WorkRows& R, int nc, int D, float* out, int lane, W w) {
    const int li = lane & 15, lk = lane >> 4;
    mfma_v4f acc[kMfmaMaxDofs / 16];
#pragma unroll
    for (int t = 0; t < kMfmaMaxDofs / 16; t++) acc[t] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < nc; c0 += 16) {
        float av[4], bv[4][kMfmaMaxDofs / 16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int c = c0 + 4 * q + lk;
            av[q] = c < nc ? w(c) : 0.f;
#pragma unroll
            for (int t = 0; t < kMfmaMaxDofs / 16; t++) {
                const int d = t * 16 + li;
                bv[q][t] = (c < nc && d < D) ? R.jac(c, d) : 0.f;
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int t = 0; t < kMfmaMaxDofs / 16; t++)
                if (t * 16 < D) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[q], bv[q][t], acc[t], 0, 0, 0);
    }
    if (lk == 0) {
#pragma unroll
        for (int t = 0; t < kMfmaMaxDofs / 16; t++)
            if (t * 16 + li < D) out[t * 16 + li] = acc[t][0];
    }
    wave_sync();
}

// R.jv(c) = sum over d of J[c][d] v[d] for c < nc (D <= 128) on the matrix cores: 16-row tiles, four at a time, A the
// rows (A[i][k] = J[c0 + i][d0 + k]) and B the vector over the tile's columns (B[k][j] = v[d0 + k]), so every column
// of a tile's result is the row sums.
__device__ inline void gemv_rows_mfma(const WorkRows& R, int nc, int D, const float* v, int lane) {
    const int li = lane & 15, lk = lane >> 4;
    for (int c0 = 0; c0 < nc; c0 += 64) {
        mfma_v4f acc[4];
#pragma unroll
        for (int q = 0; q < 4; q++) acc[q] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
        for (int d0 = 0; d0 < D; d0 += 4) {
            const int d = d0 + lk;
            const float b = d < D ? v[d] : 0.f;
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int c = c0 + 16 * q + li;
                a[q] = (c < nc && d < D) ? R.jac(c, d) : 0.f;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[q], b, acc[q], 0, 0, 0);
        }
        if (li == 0) {
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int c = c0 + 16 * q + 4 * lk + r;
                    if (c < nc) R.jv(c) = acc[q][r];
                }
        }
    }
    wave_sync();
}

// solver.py:2620-2694 (same arithmetic as dev_update_constraint; for the matrix-core models J^T force on the matrix
// cores and the cost's row sums reduced across the wavefront)
__device__ inline float wave_update_constraint(const DModel& m, const LState& loc, const WorkRows& R, int lane,
                                               float cost_in) {
    const int D = m.D;
    const int nc = LF(n_con, 0), ne = LF(n_con_eq, 0), nef = ne + LF(n_con_fl, 0);
    LF(prev_cost, 0) = cost_in;
    for (int c = lane; c < nc; c += kWave) {
        R.prev(c) = R.active(c);
        int act = 1;
        float ff = 0.f, fc = 0.f;
        float ja = R.Jaref(c);
        if (ne <= c && c < nef) {
            float f = R.fl(c), r = R.diag(c);
            float rf = r * f;
            bool ln = ja <= -rf, lp = ja >= rf;
            act = !(ln || lp);
            ff = (float)ln * f + (float)lp * -f;
            fc = (float)ln * f * (-0.5f * rf - ja);
            fc = fc + (float)lp * f * (-0.5f * rf + ja);
        } else if (nef <= c) {
            act = ja < 0.f;
        }
        R.active(c) = act;
        float Dc = R.efc_D(c);
        R.force(c) = ff + (-ja * Dc * (float)act);
        R.tmp(0, c) = fc;
        R.tmp(1, c) = 0.5f * (ja * ja * Dc * (float)act);
    }
    wave_sync();
    if (mfma_model(D)) {
        gemv_t_mfma(R, nc, D, FP(loc, qfrc_constraint), lane, [&](int c) { return R.force(c); });
        float pf = 0.f, pq = 0.f;
        for (int c = lane; c < nc; c += kWave) {
            if (ne <= c && c < nef) pf = pf + R.tmp(0, c);
            pq = pq + R.tmp(1, c);
        }
        float cost = wave_sum_f(pf);
        const float gauss = wave_dsum(D, lane, [&](int d) {
            return 0.5f * (LF(Ma, d) - LF(force, d)) * (LF(qacc, d) - LF(acc_smooth, d));
        });
        cost = cost + gauss;
        cost = cost + wave_sum_f(pq);
        LF(gauss, 0) = gauss;
        wave_sync();
        return cost;
    }
    for (int d = lane; d < D; d += kWave)
        LF(qfrc_constraint, d) = seq_sum(0.f, 0, nc, [&](int c) { return R.jac(c, d) * R.force(c); });
    // uniform: the cost in the reference's order (friction-loss terms, Gauss terms, quadratic row terms)
    float cost = seq_sum(0.f, ne, nef, [&](int c) { return R.tmp(0, c); });
    float gauss = 0.f;
    for (int d = 0; d < D; d++) {
        float v = 0.5f * (LF(Ma, d) - LF(force, d)) * (LF(qacc, d) - LF(acc_smooth, d));
        gauss = gauss + v;
        cost = cost + v;
    }
    cost = seq_sum(cost, 0, nc, [&](int c) { return R.tmp(1, c); });
    LF(gauss, 0) = gauss;
    wave_sync();
    return cost;
}

// solver.py:1595-1621 func_cholesky_factor_direct_batch on the lower triangle of the D x D matrix H (row-major), in
// place: column i's diagonal first (one chain on every lane), then the entries below it one lane per row, each sum
// over k taken in order. The terms' loads are issued 8 at a time ahead of their adds, since H may live in global
// memory.
__device__ inline void cholesky_in_place(int D, float* H, float eps, int lane) {
    constexpr int U = 8;
    for (int i = 0; i < D; i++) {
        const float* Hi = H + (size_t)i * D;
        float t = Hi[i];
        int j = 0;
        for (; j + U <= i; j += U) {
            float v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = Hi[j + u];
#pragma unroll
            for (int u = 0; u < U; u++) t = t - v[u] * v[u];
        }
        for (; j < i; j++) t = t - Hi[j] * Hi[j];
        const float dii = sqrtf(fmaxf(t, eps));
        const float inv = 1.f / dii;
        for (int r = i + 1 + lane; r < D; r += kWave) {
            const float* Hr = H + (size_t)r * D;
            float dt = 0.f;
            int k = 0;
            for (; k + U <= i; k += U) {
                float a[U], c[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    a[u] = Hr[k + u];
                    c[u] = Hi[k + u];
                }
#pragma unroll
                for (int u = 0; u < U; u++) dt = dt + a[u] * c[u];
            }
            for (; k < i; k++) dt = dt + Hr[k] * Hi[k];
            H[(size_t)r * D + i] = (Hr[i] - dt) * inv;
        }
        wave_sync();
        H[(size_t)i * D + i] = dii;
        wave_sync();
    }
}

// solver.py:1395-1452 (func_hessian_direct_tiled's product J^T diag(efc_D active) J) on the matrix cores, for models
// above the row-tile classes (dense rows over many dofs): the lower triangle in 16 x 16 tiles, all accumulated
// together over the rows 4 at a time with v_mfma_f32_16x16x4f32 (A: lane l holds J[c0 + l/16][I*16 + l%16]; B: lane l holds
// efc_D active J[c0 + l/16][J*16 + l%16]; tile element (4 (l/16) + r, l%16) in lane l's r-th accumulator). As in the
// reference, a row's term of entry (i, j) is left out when |J[c][i]| <= EPS; the mass matrix is added afterwards on
// blocks of one entity. The matrix cores sum the rows in their own order, so the entries differ from the
// one-lane-per-entry sum (and from the per-stage kernels) in the last bits.
__device__ inline void wave_hessian_mfma(const DModel& m, const LState& loc, const WorkRows& R, int lane) {
    constexpr int kT = kMfmaMaxDofs / 16;
    constexpr int kTiles = kT * (kT + 1) / 2;  // the lower triangle's 16 x 16 tiles, (I, J <= I) in row order
    const int D = m.D;
    const int nc = LF(n_con, 0);
    const int T = (D + 15) / 16;
    const float eps = m.eps;
    const int li = lane & 15, lk = lane >> 4;
    // A (lane: row i = li of tile row I, k = lk) and B (lane: k = lk, column j = li of tile column J) read the same
    // element J[c0 + lk][16 I + li], so each 4-row step loads every row's 16-column slices once and feeds all the
    // tiles from them; two steps per trip keep 16 loads in flight.
    mfma_v4f acc[kTiles];
#pragma unroll
    for (int t = 0; t < kTiles; t++) acc[t] = mfma_v4f{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < nc; c0 += 8) {
        float v[2][kT], w[2];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int c = c0 + 4 * s + lk;
            const bool rv = c < nc;
            w[s] = rv ? R.efc_D(c) * (float)R.active(c) : 0.f;
#pragma unroll
            for (int I = 0; I < kT; I++) {
                const int d = I * 16 + li;
                v[s][I] = (rv && d < D) ? R.jac(c, d) : 0.f;
            }
        }
#pragma unroll
        for (int s = 0; s < 2; s++) {
            float av[kT], bv[kT];
#pragma unroll
            for (int I = 0; I < kT; I++) {
                av[I] = fabsf(v[s][I]) > eps ? v[s][I] : 0.f;
                bv[I] = v[s][I] * w[s];
            }
#pragma unroll
            for (int I = 0; I < kT; I++)
#pragma unroll
                for (int J = 0; J <= I; J++)
                    if (I < T) acc[I * (I + 1) / 2 + J] =
                        __builtin_amdgcn_mfma_f32_16x16x4f32(av[I], bv[J], acc[I * (I + 1) / 2 + J], 0, 0, 0);
        }
    }
#pragma unroll
    for (int I = 0; I < kT; I++) {
#pragma unroll
        for (int J = 0; J <= I; J++) {
            if (I >= T) continue;
            const int j = J * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = I * 16 + 4 * lk + r;
                if (i < D && j <= i) {
                    float h = acc[I * (I + 1) / 2 + J][r];
                    if (m.dof_entity[i] == m.dof_entity[j]) h = h + LF(mass_mat, i * D + j);
                    LF(nt_H, i * D + j) = h;
                }
            }
        }
    }
    wave_sync();
}

// solver.py:1595-1621 in 16-column panels, for the models whose Hessian is assembled on the matrix cores: each panel's
// columns are factored as in cholesky_in_place for every row at or below the panel, with the sums over k starting at
// the panel's first column (the earlier panels' terms were subtracted already), then the panel's product
// L[rows][panel] L[cols][panel]^T is subtracted from the trailing lower triangle in 16 x 16 tiles on the matrix cores.
// The same factor as cholesky_in_place up to the order of the sums. H: D x D row-major, lower triangle, D <= 128.
__device__ inline void cholesky_blocked_mfma(int D, float* H, float eps, int lane) {
    const int T = (D + 15) / 16;
    const int li = lane & 15, lk = lane >> 4;
    for (int p = 0; p < T; p++) {
        const int k0 = p * 16, k1 = min(k0 + 16, D);
        for (int i = k0; i < k1; i++) {
            const float* Hi = H + (size_t)i * D;
            float hv[16];
#pragma unroll
            for (int u = 0; u < 16; u++) hv[u] = k0 + u < i ? Hi[k0 + u] : 0.f;
            float t = Hi[i];
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (k0 + u < i) t = t - hv[u] * hv[u];
            const float dii = sqrtf(fmaxf(t, eps));
            const float inv = 1.f / dii;
            for (int r = i + 1 + lane; r < D; r += kWave) {
                const float* Hr = H + (size_t)r * D;
                float rv[16];
#pragma unroll
                for (int u = 0; u < 16; u++) rv[u] = k0 + u < i ? Hr[k0 + u] : 0.f;
                float dt = 0.f;
#pragma unroll
                for (int u = 0; u < 16; u++)
                    if (k0 + u < i) dt = dt + rv[u] * hv[u];
                H[(size_t)r * D + i] = (Hr[i] - dt) * inv;
            }
            wave_sync();
            H[(size_t)i * D + i] = dii;
            wave_sync();
        }
        for (int Rt = p + 1; Rt < T; Rt++) {
            for (int Ct = p + 1; Ct <= Rt; Ct++) {
                const int ra = Rt * 16 + li, rb = Ct * 16 + li;
                float a[4], b[4], h[4];
#pragma unroll
                for (int kk = 0; kk < 4; kk++) {
                    const int k = k0 + 4 * kk + lk;
                    a[kk] = (ra < D && k < k1) ? H[(size_t)ra * D + k] : 0.f;
                    b[kk] = (rb < D && k < k1) ? H[(size_t)rb * D + k] : 0.f;
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = Rt * 16 + 4 * lk + r;
                    h[r] = (i < D && rb <= i) ? H[(size_t)i * D + rb] : 0.f;
                }
                mfma_v4f acc = mfma_v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kk = 0; kk < 4; kk++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[kk], b[kk], acc, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int i = Rt * 16 + 4 * lk + r;
                    if (i < D && rb <= i) H[(size_t)i * D + rb] = h[r] - acc[r];
                }
            }
        }
        wave_sync();
    }
}

// solver.py:1395-1452 + :1595-1621: H = M + J^T diag(D active) J (lower), then in-place Cholesky
__device__ inline void wave_hessian_cholesky(const DModel& m, const LState& loc, const WorkRows& R, int lane) {
    const int D = m.D;
    if (mfma_model(D)) {
        wave_hessian_mfma(m, loc, R, lane);
        cholesky_blocked_mfma(D, FP(loc, nt_H), m.eps, lane);
        return;
    }
    const int nc = LF(n_con, 0);
    const int nE = D * (D + 1) / 2;
    const float eps = m.eps;
    for (int e = lane; e < nE; e += kWave) {
        int i = (int)((sqrtf(8.f * (float)e + 1.f) - 1.f) * 0.5f);
        while ((i + 1) * (i + 2) / 2 <= e) i++;
        while (i * (i + 1) / 2 > e) i--;
        const int j = e - i * (i + 1) / 2;
        float h = 0.f;
        constexpr int U = 8;
        int c = 0;
        for (; c + U <= nc; c += U) {
            float v[U];
            bool use[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                float ji = R.jac(c + u, i);
                use[u] = fabsf(ji) > eps;
                v[u] = R.jac(c + u, j) * ji * R.efc_D(c + u) * (float)R.active(c + u);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (use[u]) h = h + v[u];
        }
        for (; c < nc; c++) {
            float ji = R.jac(c, i);
            if (fabsf(ji) > eps) h = h + R.jac(c, j) * ji * R.efc_D(c) * (float)R.active(c);
        }
        if (m.dof_entity[i] == m.dof_entity[j]) h = h + LF(mass_mat, i * D + j);
        LF(nt_H, i * D + j) = h;
    }
    wave_sync();
    cholesky_in_place(D, FP(loc, nt_H), eps, lane);
}

// solver.py:1782-1822: rank-1 updates of the factor for rows whose active flag flipped, in row order;
// returns true (uniformly) when an update degenerates
__device__ inline bool wave_cholesky_incremental(const DModel& m, const LState& loc, const WorkRows& R, int lane) {
    const int D = m.D;
    const int nc = LF(n_con, 0);
    for (int c0 = 0; c0 < nc; c0 += kWave) {
        const int cl = c0 + lane;
        uint64_t flips = wave_ballot(cl < nc && (R.active(cl) ^ R.prev(cl)));
        while (flips) {
            const int c = c0 + __builtin_ctzll(flips);
            flips &= flips - 1;
            const float sign = R.active(c) ? 1.f : -1.f;
            const float sq = sqrtf(R.efc_D(c));
            for (int d = lane; d < D; d += kWave) LF(nt_vec, d) = R.jac(c, d) * sq;
            wave_sync();
            for (int k = 0; k < D; k++) {
                const float vk = LF(nt_vec, k);
                if (!(fabsf(vk) > m.eps)) continue;
                const float Lkk = LF(nt_H, k * D + k);
                const float t = Lkk * Lkk + sign * vk * vk;
                if (t < m.eps) return true;
                const float r = sqrtf(t);
                const float cc = r / Lkk, cinv = 1.f / cc, ss = vk / Lkk;
                for (int i = k + 1 + lane; i < D; i += kWave) {
                    float h = (LF(nt_H, i * D + k) + ss * LF(nt_vec, i) * sign) * cinv;
                    LF(nt_H, i * D + k) = h;
                    LF(nt_vec, i) = LF(nt_vec, i) * cc - ss * h;
                }
                wave_sync();
                LF(nt_H, k * D + k) = r;
            }
        }
    }
    wave_sync();
    return false;
}

// solver.py:2722-2751 + :1893-1910 (uniform: D is small; the order of every sum is the reference's)
__device__ inline void wave_gradient(const DModel& m, const LState& loc) {
    const int D = m.D;
    for (int d = 0; d < D; d++) LF(grad, d) = LF(Ma, d) - LF(force, d) - LF(qfrc_constraint, d);
    for (int i = 0; i < D; i++) {
        float cur = LF(grad, i);
        for (int j = 0; j < i; j++) cur = cur - LF(nt_H, i * D + j) * LF(Mgrad, j);
        LF(Mgrad, i) = cur / LF(nt_H, i * D + i);
    }
    for (int ii = 0; ii < D; ii++) {
        int i = D - 1 - ii;
        float cur = LF(Mgrad, i);
        for (int j = i + 1; j < D; j++) cur = cur - LF(nt_H, j * D + i) * LF(Mgrad, j);
        LF(Mgrad, i) = cur / LF(nt_H, i * D + i);
    }
}

// solver.py:2722-2751 with the conjugate-gradient solver's preconditioner, Mgrad = M^-1 grad (solver.py:2737-2747,
// func_solve_mass_batch; forward_dynamics.py:813-894): the gradient uniform across the wave, the solve one entity per
// lane (the same arithmetic per entity)
__device__ inline void wave_gradient_cg(const DModel& m, const LState& loc, int lane) {
    const int D = m.D;
    for (int d = 0; d < D; d++) LF(grad, d) = LF(Ma, d) - LF(force, d) - LF(qfrc_constraint, d);
    wave_solve_mass(m, loc, lane, FP(loc, grad), FP(loc, Mgrad));
}

// solver.py:2722-2751 + :1893-1910 for models with 25 to 128 dofs (whose factor comes from the matrix cores),
// lane-parallel with lane l holding entries l and l + 64, reading the factor by rows only (a row of the row-major
// lower triangle is contiguous, so each load is coalesced; a column would be 64 scattered rows). Forward substitution
// row by row: y_i = (g_i - sum over k < i of L[i][k] y_k) / L[i][i], the sum formed by lanes k and reduced across the
// wavefront. Backward substitution column by column: x_i = y_i / L[i][i], then every lane k < i subtracts L[i][k] x_i
// from its y_k (row i again). The sums run in their own order (a tree forward, descending i backward), as the
// matrix-core factor already does. Rows are loaded kAhead steps early (the factor may live in global memory).
__device__ inline void wave_gradient_lanes(const DModel& m, const LState& loc, int lane) {
    constexpr int kAhead = 4;
    const int D = m.D;
    const float* L = FP(loc, nt_H);
    float cur[2], hd[2], y[2] = {0.f, 0.f}, x[2] = {0.f, 0.f};
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int i = lane + r * kWave;
        cur[r] = 0.f;
        hd[r] = 1.f;
        if (i < D) {
            const float g = LF(Ma, i) - LF(force, i) - LF(qfrc_constraint, i);
            LF(grad, i) = g;
            cur[r] = g;
            hd[r] = L[i * D + i];
        }
    }
    // L[i][k] for k < i (row i, strictly below the diagonal), else 0
    auto rowv = [&](int i, int k) { return (i >= 0 && i < D && k < i) ? L[i * D + k] : 0.f; };
    float pf[kAhead][2];
#pragma unroll
    for (int k = 0; k < kAhead; k++)
#pragma unroll
        for (int r = 0; r < 2; r++) pf[k][r] = rowv(k, lane + r * kWave);
    for (int i0 = 0; i0 < D; i0 += kAhead) {
#pragma unroll
        for (int k = 0; k < kAhead; k++) {
            const int i = i0 + k;
            if (i < D) {
                const float prod = pf[k][0] * y[0] + pf[k][1] * y[1];  // y_k is zero until set, L[i][k] zero for k >= i
#pragma unroll
                for (int r = 0; r < 2; r++) pf[k][r] = rowv(i + kAhead, lane + r * kWave);
                const int ri = i >> 6, li = i & (kWave - 1);
                float c = readlane_f(ri ? cur[1] : cur[0], li) - wave_sum_f(prod);
                c = c / readlane_f(ri ? hd[1] : hd[0], li);
                if (lane == li) {
                    if (ri) y[1] = c;
                    else y[0] = c;
                }
            }
        }
    }
    float pb[kAhead][2];
#pragma unroll
    for (int k = 0; k < kAhead; k++)
#pragma unroll
        for (int r = 0; r < 2; r++) pb[k][r] = rowv(D - 1 - k, lane + r * kWave);
    for (int ii0 = 0; ii0 < D; ii0 += kAhead) {
#pragma unroll
        for (int k = 0; k < kAhead; k++) {
            const int i = D - 1 - (ii0 + k);
            if (i >= 0) {
                const int ri = i >> 6, li = i & (kWave - 1);
                const float xi = readlane_f(ri ? y[1] : y[0], li) / readlane_f(ri ? hd[1] : hd[0], li);
                if (lane == li) {
                    if (ri) x[1] = xi;
                    else x[0] = xi;
                }
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    y[r] = y[r] - pb[k][r] * xi;  // L[i][k] is zero for lanes k >= i
                    pb[k][r] = rowv(i - kAhead, lane + r * kWave);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
        if (lane + r * kWave < D) LF(Mgrad, lane + r * kWave) = x[r];
    wave_sync();
}


// ------------------------------------------------------------------------------------------------------------
// Row-tile variants. Every fused launch gives each wavefront a row tile in LDS; each lane computes its row's terms
// in registers and stages them in the tile, and every sum over rows runs in one lane per output, in row order
// (tile_seq_sums), so many ordered sums proceed side by side. The line-search passes serve every model (KD = 0:
// any number of dofs); the constraint update, Hessian, rank-1 update and triangular-solve variants need at most KD
// dofs (instantiated for KD = 16 and KD = 24). Same arithmetic and order as the functions above.

// solver.py:2620-2694
template <int KD>
__device__ inline float wave_update_constraint_r(const DModel& m, const LState& loc, const WorkRows& R, int lane,
                                                 float cost_in, float* tile) {
    const int D = m.D;
    const int nc = LF(n_con, 0), ne = LF(n_con_eq, 0), nef = ne + LF(n_con_fl, 0);
    LF(prev_cost, 0) = cost_in;
    // friction-loss terms first (the reference adds them to the cost inside its row loop, before anything else)
    float cost[1] = {0.f};
    for (int c = ne; c < nef; c++) {
        const float f = R.fl(c), r = R.diag(c), ja = R.Jaref(c);
        const float rf = r * f;
        const bool ln = ja <= -rf, lp = ja >= rf;
        float fc = (float)ln * f * (-0.5f * rf - ja);
        fc = fc + (float)lp * f * (-0.5f * rf + ja);
        cost[0] = cost[0] + fc;
    }
    float gauss = 0.f;
    for (int d = 0; d < D; d++) {
        float v = 0.5f * (LF(Ma, d) - LF(force, d)) * (LF(qacc, d) - LF(acc_smooth, d));
        gauss = gauss + v;
        cost[0] = cost[0] + v;
    }
    // lane d < D sums qfrc_constraint[d] = sum_c jac[c][d] force[c] (tile rows d and D + 1), lane D the cost's
    // quadratic terms (tile row D, times the ones of row D + 2, which leaves them exact), both in row order
    float q = lane == D ? cost[0] : 0.f;
    tile[(D + 2) * kTileStride + lane] = 1.f;
    // this lane's row of the next chunk is loaded while the current chunk is summed
    float n_ja = 0.f, n_f = 0.f, n_r = 0.f, n_D = 0.f, n_j[KD];
    int n_act = 0;
    auto load = [&](int c) {
        if (c < nc) {
            n_act = R.active(c);
            n_ja = R.Jaref(c);
            n_D = R.efc_D(c);
            if (ne <= c && c < nef) {
                n_f = R.fl(c);
                n_r = R.diag(c);
            }
#pragma unroll
            for (int d = 0; d < KD; d++)
                if (d < D) n_j[d] = R.jac(c, d);
        }
    };
    load(lane);
    for (int c0 = 0; c0 < nc; c0 += kWave) {
        const int c = c0 + lane;
        const int prev_act = n_act;
        const float ja = n_ja, fl = n_f, dg = n_r, Dc = n_D;
        float jr[KD];
#pragma unroll
        for (int d = 0; d < KD; d++) jr[d] = n_j[d];
        load(c + kWave);
        if (c < nc) {
            R.prev(c) = prev_act;
            int act = 1;
            float ff = 0.f;
            if (ne <= c && c < nef) {
                const float f = fl, r = dg;
                const float rf = r * f;
                const bool ln = ja <= -rf, lp = ja >= rf;
                act = !(ln || lp);
                ff = (float)ln * f + (float)lp * -f;
            } else if (nef <= c) {
                act = ja < 0.f;
            }
            R.active(c) = act;
            const float force = ff + (-ja * Dc * (float)act);
            R.force(c) = force;
#pragma unroll
            for (int d = 0; d < KD; d++)
                if (d < D) tile[d * kTileStride + lane] = jr[d];
            tile[D * kTileStride + lane] = 0.5f * (ja * ja * Dc * (float)act);
            tile[(D + 1) * kTileStride + lane] = force;
        }
        const int n = min(kWave, nc - c0);
        sync_for(tile);
        if (lane <= D) {
            const float* col = tile + lane * kTileStride;
            const float* fc = tile + (lane < D ? D + 1 : D + 2) * kTileStride;
            if (n == kWave) q = row_dot64(q, col, fc);
            else q = seq_sum<16>(q, 0, n, [&](int l) { return col[l] * fc[l]; });
        }
        sync_for(tile);
    }
    if (lane < D) LF(qfrc_constraint, lane) = q;
    LF(gauss, 0) = gauss;
    cost[0] = readlane_f(q, D);
    wave_sync();
    return cost[0];
}

// solver.py:1395-1452 + :1595-1621 with each chunk of rows staged in the row tile: one lane per Hessian entry sums
// its terms over the rows in row order from LDS
template <int KD>
__device__ inline void wave_hessian_cholesky_tiled(const DModel& m, const LState& loc, const WorkRows& R, int lane,
                                                   float* tile) {
    const int D = m.D;
    const int nc = LF(n_con, 0);
    const int nE = D * (D + 1) / 2;
    const float eps = m.eps;
    constexpr int kE = (KD * (KD + 1) / 2 + kWave - 1) / kWave;  // entries per lane
    float h[kE];
    int ei[kE], ej[kE];
#pragma unroll
    for (int r = 0; r < kE; r++) {
        h[r] = 0.f;
        const int e = lane + r * kWave;
        int i = (int)((sqrtf(8.f * (float)e + 1.f) - 1.f) * 0.5f);
        while ((i + 1) * (i + 2) / 2 <= e) i++;
        while (i * (i + 1) / 2 > e) i--;
        ei[r] = i;
        ej[r] = e - i * (i + 1) / 2;
    }
    // this lane's row of the next chunk is loaded while the current chunk is summed
    float n_j[KD], n_D = 0.f;
    int n_act = 0;
    auto load = [&](int c) {
        if (c < nc) {
            n_D = R.efc_D(c);
            n_act = R.active(c);
#pragma unroll
            for (int d = 0; d < KD; d++)
                if (d < D) n_j[d] = R.jac(c, d);
        }
    };
    load(lane);
    for (int c0 = 0; c0 < nc; c0 += kWave) {
        const int c = c0 + lane;
        if (c < nc) {
#pragma unroll
            for (int d = 0; d < KD; d++)
                if (d < D) tile[d * kTileStride + lane] = n_j[d];
            tile[D * kTileStride + lane] = n_D;
            tile[(D + 1) * kTileStride + lane] = (float)n_act;
        }
        load(c + kWave);
        const int n = min(kWave, nc - c0);
        sync_for(tile);
#pragma unroll
        for (int r = 0; r < kE; r++) {
            if (lane + r * kWave < nE) {
                const float* ci = tile + ei[r] * kTileStride;
                const float* cj = tile + ej[r] * kTileStride;
                const float* cd = tile + D * kTileStride;
                const float* ca = tile + (D + 1) * kTileStride;
                float acc = h[r];
                // terms in row order; the loads of 8 rows (two 16-byte loads per tile row) come before their adds
                int l = 0;
                for (; l + 8 <= n; l += 8) {
                    float t[8];
                    bool use[8];
#pragma unroll
                    for (int hh = 0; hh < 2; hh++) {
                        const float4 vi = ((const float4*)(ci + l))[hh], vj = ((const float4*)(cj + l))[hh];
                        const float4 vd = ((const float4*)(cd + l))[hh], va = ((const float4*)(ca + l))[hh];
                        const float ji[4] = {vi.x, vi.y, vi.z, vi.w}, jj[4] = {vj.x, vj.y, vj.z, vj.w};
                        const float dd[4] = {vd.x, vd.y, vd.z, vd.w}, aa[4] = {va.x, va.y, va.z, va.w};
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            use[4 * hh + u] = fabsf(ji[u]) > eps;
                            t[4 * hh + u] = jj[u] * ji[u] * dd[u] * aa[u];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++)
                        if (use[u]) acc = acc + t[u];
                }
                for (; l < n; l++) {
                    const float ji = ci[l];
                    if (fabsf(ji) > eps) acc = acc + cj[l] * ji * cd[l] * ca[l];
                }
                h[r] = acc;
            }
        }
        sync_for(tile);
    }
#pragma unroll
    for (int r = 0; r < kE; r++) {
        const int e = lane + r * kWave;
        if (e < nE) {
            int i = (int)((sqrtf(8.f * (float)e + 1.f) - 1.f) * 0.5f);
            while ((i + 1) * (i + 2) / 2 <= e) i++;
            while (i * (i + 1) / 2 > e) i--;
            const int j = e - i * (i + 1) / 2;
            float v = h[r];
            if (m.dof_entity[i] == m.dof_entity[j]) v = v + LF(mass_mat, i * D + j);
            LF(nt_H, i * D + j) = v;
        }
    }
    wave_sync();
    cholesky_in_place(D, FP(loc, nt_H), eps, lane);
}

// solver.py:1782-1822 with the rank-1 updates of consecutive flipped rows pipelined. Update u (the u-th flipped row in
// row order) applies its rotation for column k at step start(u) + k, and start times strictly increase, so the updates
// in flight at one step work on different columns of the factor, while every element of the factor still receives the
// updates in row order, each with the reference's arithmetic. The lanes form S = min(64 / D, D) groups of D lanes;
// group s carries the update in slot s (lane i of the group holds element i of its working vector), and slots take
// updates round robin, at most one new update per step. Rows are scanned 64 at a time; the scaled jacobian rows
// jac[c] * sqrt(efc_D[c]) of a chunk's flipped rows are staged in the row tile, so that starting an update reads LDS
// only. An update that degenerates makes the whole call return true (uniformly): the caller refactors from scratch,
// as the reference does once any update has degenerated.
template <int KD>
__device__ inline bool wave_cholesky_incremental_r(const DModel& m, const LState& loc, const WorkRows& R, int lane,
                                                   float* tile) {
    const int D = m.D;
    const int nc = LF(n_con, 0);
    if (D <= 0) return false;
    const int S = min(kWave / D, D);
    const int s = lane / D, i = lane - s * D;
    const bool inslot = s < S;
    const int first = (inslot ? s : 0) * D;  // first lane of this lane's group
    int k = D;                                // column the group's update works on next; D when the slot is idle
    float sign = 1.f, v = 0.f;
    int c0 = -kWave, act = 0;
    uint64_t flips = 0;
    int next = 0, busy = 0;
    // this lane's row of the next chunk (flags, efc_D, jacobian row) is loaded one chunk ahead
    int p_act = 0, p_prev = 0;
    float p_D = 0.f, p_j[KD];
    auto fetch = [&](int cl) {
        p_act = p_prev = 0;
        if (cl < nc) {
            p_act = R.active(cl);
            p_prev = R.prev(cl);
            p_D = R.efc_D(cl);
#pragma unroll
            for (int d = 0; d < KD; d++)
                if (d < D) p_j[d] = R.jac(cl, d);
        }
    };
    fetch(lane);
    for (;;) {
        while (!flips && c0 + kWave < nc) {
            c0 += kWave;
            wave_sync();  // earlier readers of the tile are done
            act = p_act;
            const bool flip = (p_act ^ p_prev) != 0;
            if (flip) {
                const float sq = sqrtf(p_D);
#pragma unroll
                for (int d = 0; d < KD; d++)
                    if (d < D) tile[d * kTileStride + lane] = p_j[d] * sq;
            }
            flips = wave_ballot(flip);
            fetch(c0 + kWave + lane);
            wave_sync();
        }
        if (flips && busy < S) {  // slots finish in the order they start: slot `next` is the idle one
            const int l = __builtin_ctzll(flips);
            flips &= flips - 1;
            const float sg = __builtin_amdgcn_readlane(act, l) ? 1.f : -1.f;
            if (s == next) {
                k = 0;
                sign = sg;
                v = tile[i * kTileStride + l];
            }
            next = next + 1 == S ? 0 : next + 1;
        }
        const bool run = inslot && k < D;
        if (!wave_ballot(run)) break;  // nothing in flight and no flipped row left
        const int kk = run ? k : 0;
        const float vk = __int_as_float(__builtin_amdgcn_ds_bpermute((first + kk) << 2, __float_as_int(v)));
        const float Lkk = LF(nt_H, kk * D + kk);
        const bool below = run && i > kk;
        const float hik = below ? LF(nt_H, i * D + kk) : 0.f;
        const bool rot = run && fabsf(vk) > m.eps;
        const float t = Lkk * Lkk + sign * vk * vk;
        if (wave_ballot(rot && t < m.eps)) return true;
        const float r = sqrtf(t);
        const float cc = r / Lkk, cinv = 1.f / cc, ss = vk / Lkk;
        if (rot && below) {
            const float h = (hik + ss * v * sign) * cinv;
            LF(nt_H, i * D + kk) = h;
            v = v * cc - ss * h;
        }
        if (rot && i == kk) LF(nt_H, kk * D + kk) = r;
        if (run) k++;
        busy = __popcll(wave_ballot(inslot && i == 0 && k < D));
        sync_for(FP(loc, nt_H));
    }
    wave_sync();
    return false;
}

// solver.py:2722-2751 + :1893-1910. Forward substitution column by column (lane i subtracts L[i][j] x_j for j = 0, 1,
// ... in turn - the reference's order for row i); backward substitution row by row in the reference's order, with
// x_j and L[j][i] read from lane j.
template <int KD>
__device__ inline void wave_gradient_r(const DModel& m, const LState& loc, int lane) {
    const int D = m.D;
    const bool own = lane < D;
    float cur = 0.f, hd = 1.f, x = 0.f;
    if (own) {
        const float g = LF(Ma, lane) - LF(force, lane) - LF(qfrc_constraint, lane);
        LF(grad, lane) = g;
        cur = g;
        hd = LF(nt_H, lane * D + lane);
    }
    for (int j = 0; j < D; j++) {
        const float hij = (own && lane > j) ? LF(nt_H, lane * D + j) : 0.f;
        const float xj = readlane_f(cur, j) / readlane_f(hd, j);
        if (lane == j) x = xj;
        if (own && lane > j) cur = cur - hij * xj;
    }
    for (int ii = 0; ii < D; ii++) {
        const int i = D - 1 - ii;
        const float col = (own && lane > i) ? LF(nt_H, lane * D + i) : 0.f;  // lane j holds L[j][i]
        float c = readlane_f(x, i);
        for (int j = i + 1; j < D; j++) c = c - readlane_f(col, j) * readlane_f(x, j);
        c = c / readlane_f(hd, i);
        if (lane == i) x = c;
    }
    if (own) LF(Mgrad, lane) = x;
    wave_sync();
}

// solver.py:2034-2152
template <int KD>
__device__ inline LsP wave_ls_init_r(const DModel& m, const LState& loc, const WorkRows& R, int lane, int& ls_it,
                                     float* tile) {
    const int D = m.D;
    const int nc = LF(n_con, 0), ne = LF(n_con_eq, 0), nef = ne + LF(n_con_fl, 0);
    for (int i = lane; i < D; i += kWave) {
        const int e = m.dof_entity[i];
        float mv = 0.f;
        for (int j = m.entity_dof_start[e]; j < m.entity_dof_end[e]; j++) mv = mv + LF(mass_mat, i * D + j) * LF(search, j);
        LF(mv, i) = mv;
    }
    constexpr int KR = KD > 0 ? KD : 1;  // register copies of the search direction and of a jacobian row (KD > 0)
    float sr[KR];
#pragma unroll
    for (int d = 0; d < KR; d++) sr[d] = (KD > 0 && d < D) ? LF(search, d) : 0.f;
    wave_sync();
    float g1 = 0.f, g2 = 0.f;
    if (KD == 0 && mfma_model(D)) {
        g1 = wave_dsum(D, lane, [&](int d) {
            const float s = LF(search, d);
            return s * LF(Ma, d) - s * LF(force, d);
        });
        g2 = wave_dsum(D, lane, [&](int d) { return 0.5f * LF(search, d) * LF(mv, d); });
    } else {
        for (int d = 0; d < D; d++) {
            float s = LF(search, d);
            g1 = g1 + (s * LF(Ma, d) - s * LF(force, d));
            g2 = g2 + 0.5f * s * LF(mv, d);
        }
    }
    const float gauss = LF(gauss, 0);
    LF(quad_gauss, 0) = gauss;
    LF(quad_gauss, 1) = g1;
    LF(quad_gauss, 2) = g2;
    // the matrix-core models take J search from the matrix cores
    const bool jv_mfma = KD == 0 && mfma_model(D);
    if (jv_mfma) gemv_rows_mfma(R, nc, D, FP(loc, search), lane);
    // lanes 0-2 sum t[q] = (gauss, g1, g2)[q] + f[q] over all rows, lanes 3-5 e[q] over the equality rows, in row
    // order, from tile row q
    float acc = lane == 0 ? gauss : lane == 1 ? g1 : lane == 2 ? g2 : 0.f;
    // this lane's row of the next chunk is loaded while the current chunk is summed
    float n_ja = 0.f, n_f = 0.f, n_r = 0.f, n_D = 0.f, n_j[KR];
    auto load = [&](int c) {
        if (c < nc) {
            n_ja = R.Jaref(c);
            n_D = R.efc_D(c);
            if (ne <= c && c < nef) {
                n_f = R.fl(c);
                n_r = R.diag(c);
            }
#pragma unroll
            for (int d = 0; d < KR; d++)
                if (KD > 0 && d < D) n_j[d] = R.jac(c, d);
        }
    };
    load(lane);
    for (int c0 = 0; c0 < nc; c0 += kWave) {
        const int c = c0 + lane;
        const float ja = n_ja, Dc = n_D, fl_c = n_f, dg = n_r;
        float jr[KR];
#pragma unroll
        for (int d = 0; d < KR; d++) jr[d] = n_j[d];
        load(c + kWave);
        if (c < nc) {
            float jv = 0.f;
            if constexpr (KD > 0) {
#pragma unroll
                for (int d = 0; d < KD; d++)
                    if (d < D) jv = jv + jr[d] * sr[d];
            } else if (jv_mfma) {
                jv = R.jv(c);
            } else {
                for (int d = 0; d < D; d++) jv = jv + R.jac(c, d) * LF(search, d);
            }
            R.jv(c) = jv;
            float f0 = Dc * (0.5f * ja * ja), f1 = Dc * (jv * ja), f2 = Dc * (0.5f * jv * jv);
            if (c < ne) {
            } else if (c < nef) {
                const float fl = fl_c, r = dg, rf = r * fl;
                const bool ln = ja <= -rf, lp = ja >= rf;
                if (ln || lp) {
                    f0 = (float)ln * fl * (-0.5f * rf - ja) + (float)lp * fl * (-0.5f * rf + ja);
                    f1 = (float)ln * (-fl * jv) + (float)lp * (fl * jv);
                    f2 = 0.f;
                }
            } else {
                const float a = (float)(ja < 0.f);
                f0 = f0 * a;
                f1 = f1 * a;
                f2 = f2 * a;
            }
            tile[lane] = f0;
            tile[kTileStride + lane] = f1;
            tile[2 * kTileStride + lane] = f2;
        }
        const int n = min(kWave, nc - c0);
        sync_for(tile);
        if (lane < 6) {
            const float* row = tile + (lane < 3 ? lane : lane - 3) * kTileStride;
            const int l1 = lane < 3 ? n : max(0, min(n, ne - c0));
            if (l1 == kWave) acc = row_sum64(acc, row);
            else acc = seq_sum<16>(acc, 0, l1, [&](int l) { return row[l]; });
        }
        sync_for(tile);
    }
    const float t[3] = {readlane_f(acc, 0), readlane_f(acc, 1), readlane_f(acc, 2)};
    LF(eq_sum, 0) = readlane_f(acc, 3);
    LF(eq_sum, 1) = readlane_f(acc, 4);
    LF(eq_sum, 2) = readlane_f(acc, 5);
    float h = 2.f * t[2];
    if (h <= 0.f) h = m.eps;
    ls_it = 1;
    wave_sync();
    return {0.f, t[0], t[1], h};
}

// solver.py:2155-2355 for N = 1 or 3 candidate steps
template <int N>
__device__ inline void wave_ls_eval_r(const DModel& m, const LState& loc, const WorkRows& R, int lane, const float* al,
                                      LsP* out, int& ls_it, float* tile) {
    const int nc = LF(n_con, 0), ne = LF(n_con_eq, 0), nef = ne + LF(n_con_fl, 0);
    const float b0 = LF(quad_gauss, 0) + LF(eq_sum, 0), b1 = LF(quad_gauss, 1) + LF(eq_sum, 1),
                b2 = LF(quad_gauss, 2) + LF(eq_sum, 2);
    // lane 3k + j (j < 3) sums term j of candidate k over the rows, starting from b_j
    const int lj = lane % 3;
    float acc = lj == 0 ? b0 : lj == 1 ? b1 : b2;
    float A[N];
#pragma unroll
    for (int k = 0; k < N; k++) A[k] = al[k];
    // this lane's row of the next chunk is loaded while the current chunk is summed
    auto load = [&](int c, float& ja, float& jv, float& Dc, float& f, float& r) {
        ja = jv = Dc = f = r = 0.f;
        if (c >= ne && c < nc) {
            ja = R.Jaref(c);
            jv = R.jv(c);
            Dc = R.efc_D(c);
            if (c < nef) {
                f = R.fl(c);
                r = R.diag(c);
            }
        }
    };
    int c0 = (ne / kWave) * kWave;
    float n_ja, n_jv, n_D, n_f, n_r;
    load(c0 + lane, n_ja, n_jv, n_D, n_f, n_r);
    for (; c0 < nc; c0 += kWave) {
        const int c = c0 + lane;
        const float ja = n_ja, jv = n_jv, Dc = n_D, fl = n_f, dg = n_r;
        load(c + kWave, n_ja, n_jv, n_D, n_f, n_r);
        float v[3 * N];
#pragma unroll
        for (int q = 0; q < 3 * N; q++) v[q] = 0.f;
        if (c >= ne && c < nc) {
            const float f0 = Dc * (0.5f * ja * ja), f1 = Dc * (jv * ja), f2 = Dc * (0.5f * jv * jv);
            if (c < nef) {
                const float f = fl, r = dg, rf = r * f;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const float x = ja + A[k] * jv;
                    const bool ln = x <= -rf, lp = x >= rf;
                    float a0 = f0, a1 = f1, a2 = f2;
                    if (ln || lp) {
                        a0 = (float)ln * f * (-0.5f * rf - ja) + (float)lp * f * (-0.5f * rf + ja);
                        a1 = (float)ln * (-f * jv) + (float)lp * (f * jv);
                        a2 = 0.f;
                    }
                    v[3 * k] = a0;
                    v[3 * k + 1] = a1;
                    v[3 * k + 2] = a2;
                }
            } else {
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const float a = (float)(ja + A[k] * jv < 0.f);
                    v[3 * k] = f0 * a;
                    v[3 * k + 1] = f1 * a;
                    v[3 * k + 2] = f2 * a;
                }
            }
        }
        tile_seq_sums<3 * N>(acc, v, max(0, ne - c0), min(kWave, nc - c0), tile, lane);
    }
    float T[3 * N];
#pragma unroll
    for (int q = 0; q < 3 * N; q++) T[q] = readlane_f(acc, q);
#pragma unroll
    for (int k = 0; k < N; k++) {
        const float a = A[k];
        float h = 2.f * T[3 * k + 2];
        if (h <= 0.f) h = m.eps;
        out[k] = {a, a * a * T[3 * k + 2] + a * T[3 * k + 1] + T[3 * k], 2.f * a * T[3 * k + 2] + T[3 * k + 1], h};
    }
    ls_it += N;
}

// solver.py:2427-2597 func_linesearch_batch (uniform control flow)
template <int KD>
__device__ inline float wave_linesearch(const DModel& m, const LState& loc, const WorkRows& R, int lane, float* tile) {
    const int D = m.D;
    float sn = 0.f;
    if (KD == 0 && mfma_model(D)) {
        sn = wave_dsum(D, lane, [&](int d) { return LF(search, d) * LF(search, d); });
    } else {
        for (int d = 0; d < D; d++) sn = sn + LF(search, d) * LF(search, d);
    }
    sn = sqrtf(sn);
    float gtol = m.tolerance * m.ls_tolerance * sn * (m.meaninertia * (float)max(1, D));
    LF(gtol, 0) = gtol;
    int ls_it = 0, res_code = 0;
    float res = 0.f;
    if (sn < m.eps) {
        LF(ls_result, 0) = 1;
        LF(ls_it, 0) = 0;
        return 0.f;
    }
    LsP p0;
    p0 = wave_ls_init_r<KD>(m, loc, R, lane, ls_it, tile);
    LsP p1;
    float a = p0.alpha - p0.grad / p0.hess;
    wave_ls_eval_r<1>(m, loc, R, lane, &a, &p1, ls_it, tile);
    if (p0.cost < p1.cost) p1 = p0;
    bool done = false;
    if (fabsf(p1.grad) < gtol) {
        res_code = fabsf(p1.alpha) < m.eps ? 2 : 0;
        res = p1.alpha;
        done = true;
    }
    LsP p2 = p1;
    int p2update = 0;
    if (!done) {
        float dir = (float)((p1.grad < 0.f) ? 1 : 0) * 2.f - 1.f;
        while (p1.grad * dir <= -gtol && ls_it < m.ls_iterations) {
            p2 = p1;
            p2update = 1;
            float an = p1.alpha - p1.grad / p1.hess;
            wave_ls_eval_r<1>(m, loc, R, lane, &an, &p1, ls_it, tile);
            if (fabsf(p1.grad) < gtol) {
                res = p1.alpha;
                done = true;
                break;
            }
        }
    }
    if (!done && ls_it >= m.ls_iterations) {
        res_code = 3;
        res = p1.alpha;
        done = true;
    }
    if (!done && !p2update) {
        res_code = 6;
        res = p1.alpha;
        done = true;
    }
    if (!done) {
        float al[3] = {p1.alpha - p1.grad / p1.hess, p1.alpha, (p1.alpha + p2.alpha) * 0.5f};
        while (ls_it < m.ls_iterations) {
            LsP c[3];
            wave_ls_eval_r<3>(m, loc, R, lane, al, c, ls_it, tile);
            float n1 = al[0], n2 = al[1];
            bool found = false;
            float ba = 0.f, bc = 0.f;
            for (int i = 0; i < 3; i++)
                if (fabsf(c[i].grad) < gtol && (!found || c[i].cost < bc)) {
                    ba = c[i].alpha;
                    bc = c[i].cost;
                    found = true;
                }
            if (found) {
                res = ba;
                done = true;
                break;
            }
            int b1 = bracket(p1, c, n1), b2 = bracket(p2, c, n2);
            if (b1 == 0 && b2 == 0) {
                res_code = c[2].cost < p0.cost ? 0 : 7;
                res = al[2];
                done = true;
                break;
            }
            al[0] = n1;
            al[1] = n2;
            al[2] = (p1.alpha + p2.alpha) * 0.5f;
        }
        if (!done) {
            if (p1.cost <= p2.cost && p1.cost < p0.cost) {
                res_code = 4;
                res = p1.alpha;
            } else if (p2.cost <= p1.cost && p2.cost < p0.cost) {
                res_code = 4;
                res = p2.alpha;
            } else {
                res_code = 5;
                res = 0.f;
            }
        }
    }
    LF(ls_result, 0) = res_code;
    LF(ls_it, 0) = ls_it;
    return res;
}

// solver.py:2936-3059 (init), :3063-3139 (Newton or conjugate-gradient iteration), :3164-3190 (body), :3241-3261,
// :3199-3237
template <int KD>
__device__ inline void wave_solve(const DModel& m, const LState& loc, const DState& st, int b, const WorkRows& R,
                                  int lane, float* tile) {
    const int D = m.D;
    const int nc = LF(n_con, 0);
    const bool ws = nc > 0 && LF(is_warmstart, 0);
    // models on the matrix-core path update their dof vectors lane-parallel (entries lane, lane + 64)
    const bool lane_dofs = KD == 0 && mfma_model(D);
    for (int d = lane_dofs ? lane : 0; d < D; d += lane_dofs ? kWave : 1)
        LF(qacc, d) = ws ? LF(qacc_ws, d) : LF(acc_smooth, d);
    if (lane_dofs) wave_sync();
    for (int d1 = lane; d1 < D; d1 += kWave) {
        int e = m.dof_entity[d1];
        float ma = 0.f;
        for (int d2 = m.entity_dof_start[e]; d2 < m.entity_dof_end[e]; d2++) ma = ma + LF(mass_mat, d1 * D + d2) * LF(qacc, d2);
        LF(Ma, d1) = ma;
    }
    for (int c = lane; c < nc; c += kWave) {
        R.Jaref(c) = seq_sum(-R.aref(c), 0, D, [&](int d) { return R.jac(c, d) * LF(qacc, d); });
        R.active(c) = 0;  // the first update compares against a cleared previous active set
    }
    wave_sync();
    float cost;
    if constexpr (KD > 0) cost = wave_update_constraint_r<KD>(m, loc, R, lane, LF(cost, 0), tile);
    else cost = wave_update_constraint(m, loc, R, lane, LF(cost, 0));
    LF(cost, 0) = cost;
    int improved = nc > 0;
    const bool cg = m.solver_type == RS_SOLVER_CG;
    if (cg) {
        wave_gradient_cg(m, loc, lane);
    } else {
        if constexpr (KD > 0) wave_hessian_cholesky_tiled<KD>(m, loc, R, lane, tile);
        else wave_hessian_cholesky(m, loc, R, lane);
        if constexpr (KD > 0) wave_gradient_r<KD>(m, loc, lane);
        else if (D <= 2 * kWave) wave_gradient_lanes(m, loc, lane);
        else wave_gradient(m, loc);
    }
    if (lane_dofs) {
        for (int d = lane; d < D; d += kWave) LF(search, d) = -LF(Mgrad, d);
        wave_sync();
    } else {
        for (int d = 0; d < D; d++) LF(search, d) = -LF(Mgrad, d);
    }
    int n_iter = 0;
    if (nc > 0) {
        const float tol = (m.meaninertia * (float)max(1, D)) * m.tolerance;
        for (int it = 0; it < m.iterations; it++) {
            n_iter++;
            float alpha = wave_linesearch<KD>(m, loc, R, lane, tile);
            if (fabsf(alpha) < m.eps) {
                improved = 0;
            } else {
                for (int d = lane_dofs ? lane : 0; d < D; d += lane_dofs ? kWave : 1) {
                    LF(qacc, d) = LF(qacc, d) + LF(search, d) * alpha;
                    LF(Ma, d) = LF(Ma, d) + LF(mv, d) * alpha;
                }
                for (int c = lane; c < nc; c += kWave) R.Jaref(c) = R.Jaref(c) + R.jv(c) * alpha;
                if (cg)
                    for (int d = 0; d < D; d++) {
                        FV(st, cg_prev_grad, b)[d] = LF(grad, d);
                        FV(st, cg_prev_Mgrad, b)[d] = LF(Mgrad, d);
                    }
                wave_sync();
                float prev = LF(cost, 0);
                if constexpr (KD > 0) cost = wave_update_constraint_r<KD>(m, loc, R, lane, prev, tile);
                else cost = wave_update_constraint(m, loc, R, lane, prev);
                LF(cost, 0) = cost;
                if (cg) {
                    wave_gradient_cg(m, loc, lane);
                } else {
                    bool degenerate;
                    if constexpr (KD > 0) degenerate = wave_cholesky_incremental_r<KD>(m, loc, R, lane, tile);
                    else degenerate = wave_cholesky_incremental(m, loc, R, lane);
                    if (degenerate) {
                        if constexpr (KD > 0) wave_hessian_cholesky_tiled<KD>(m, loc, R, lane, tile);
                        else wave_hessian_cholesky(m, loc, R, lane);
                    }
                    if constexpr (KD > 0) wave_gradient_r<KD>(m, loc, lane);
                    else if (D <= 2 * kWave) wave_gradient_lanes(m, loc, lane);
                    else wave_gradient(m, loc);
                }
                float gn = 0.f;
                if (lane_dofs) {
                    gn = wave_dsum(D, lane, [&](int d) { return LF(grad, d) * LF(grad, d); });
                } else {
                    for (int d = 0; d < D; d++) gn = gn + LF(grad, d) * LF(grad, d);
                }
                gn = sqrtf(gn);
                improved = gn > tol && (prev - cost) > tol;
                if (improved && !cg) {
                    if (lane_dofs) {
                        for (int d = lane; d < D; d += kWave) LF(search, d) = -LF(Mgrad, d);
                        wave_sync();
                    } else {
                        for (int d = 0; d < D; d++) LF(search, d) = -LF(Mgrad, d);
                    }
                } else if (improved) {  // solver.py:2865-2884 Polak-Ribiere direction
                    float beta = 0.f, pg_dot_pMg = 0.f;
                    for (int d = 0; d < D; d++) {
                        beta = beta + LF(grad, d) * (LF(Mgrad, d) - FV(st, cg_prev_Mgrad, b)[d]);
                        pg_dot_pMg = pg_dot_pMg + FV(st, cg_prev_Mgrad, b)[d] * FV(st, cg_prev_grad, b)[d];
                    }
                    beta = fmaxf(beta / fmaxf(m.eps, pg_dot_pMg), 0.f);
                    FV(st, cg_pg_dot_pMg, b)[0] = pg_dot_pMg;
                    FV(st, cg_beta, b)[0] = beta;
                    for (int d = 0; d < D; d++) LF(search, d) = -LF(Mgrad, d) + beta * LF(search, d);
                }
            }
            if (!improved) break;
        }
    } else {
        improved = 0;
    }
    LF(improved, 0) = improved;
    LF(n_iter, 0) = n_iter;
    bool nan_seen = false;
    for (int d = lane_dofs ? lane : 0; d < D; d += lane_dofs ? kWave : 1) {
        float qa = LF(qacc, d);
        LF(acc, d) = qa;
        LF(qf_constraint, d) = LF(qfrc_constraint, d);
        LF(force, d) = LF(qf_smooth, d) + LF(qfrc_constraint, d);
        LF(qacc_ws, d) = qa;
        nan_seen = nan_seen || isnan(qa);
    }
    if (lane_dofs ? wave_ballot(nan_seen) != 0 : nan_seen) LF(errbits, 0) = LF(errbits, 0) | RS_ERR_INVALID_FORCE_NAN;
    if (lane_dofs) wave_sync();
    LF(is_warmstart, 0) = 1;
    // the constraint forces of the rows (ConstraintState.efc_force), from the work field to the state slab
    const int n_rows = LF(n_con, 0);
    for (int c = lane; c < n_rows; c += kWave) st.efc_force[(size_t)c * st.B + b] = R.force(c);
    const int cs = LF(n_con_eq, 0) + LF(n_con_fl, 0);
    const int ncont = LF(n_contacts, 0);
    for (int i = lane; i < ncont; i += kWave) {
        f3 nrm = ld3(st.contact_normal, i, st.B, b);
        float fr = st.contact_friction[(size_t)i * st.B + b];
        f3 d1, d2, f = zero3();
        orthogonals(nrm, d1, d2);
        for (int k = 0; k < 4; k++) {
            f3 dd = (float)(2 * (k % 2) - 1) * (k < 2 ? d1 : d2);
            f = f + (dd * fr - nrm) * R.force(i * 4 + k + cs);
        }
        st3(st.contact_force, i, st.B, b, f);
    }
    wave_sync();
}

#undef LF

}  // namespace rs
