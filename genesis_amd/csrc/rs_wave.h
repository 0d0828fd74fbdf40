// Wavefront-per-environment execution for the fused substep.
//
// One 64-lane wavefront owns one environment for the whole substep. The environment's small fields (links, joints,
// dofs, geoms, the mass matrix and solver matrices, SAP buffers) are staged from the [row][env]
// planes into a private region - LDS when it fits, else an env-major slice of the work field - and written back at
// the end. Constraint rows, per-row temporaries and per-pair contact slots live in the env-major work field. Lanes
// split row, pair, contact and matrix-entry loops; code that is inherently sequential (kinematic chains, the SAP
// sweep, triangular solves) runs uniformly on every lane of the wavefront, so its results need no broadcast. Every
// sum keeps the reference's sequential order (one lane, or every lane, walks the terms in order), so the fused step
// computes the same floating-point values as the per-stage kernels.
#pragma once
#include "rs_device.h"

namespace rs {

constexpr int kWave = 64;
constexpr int kMaxStateFields = 128;

// kinds whose fields are staged into the per-environment private region
__host__ __device__ constexpr bool is_local_kind(int kind) {
    return kind == RS_N_ONE || kind == RS_N_LINKS || kind == RS_N_JOINTS || kind == RS_N_DOFS || kind == RS_N_QS ||
           kind == RS_N_GEOMS || kind == RS_N_ENTITIES || kind == RS_N_DOFS2 || kind == RS_N_SORT;
}

// Launch parameters of the fused kernel; offsets are in floats.
struct FusedArgs {
    int epb;           // environments (wavefronts) per workgroup
    int n_local;       // local rows staged per environment (the packed local planes)
    int loc_floats;    // private region per environment (n_local rounded up to 16 bytes)
    int64_t work_stride;  // floats of the work field per environment
    int off_loc, off_jac, off_rowf, off_tmp, off_pc;
    int tile_off;      // row tile in the private region (kTileStride x (D + 2) floats), -1 when D > kTileMaxDofs
    int sched;         // 1: map wavefronts to environments through the cost-ordered queues
    int loc_off[kMaxStateFields];  // private-region offset of each state field (local kinds only)
};

// Row tile: one chunk of 64 constraint rows staged column by column (column k at k * kTileStride), so that the
// lanes that sum a column over the rows read it from LDS; the odd stride spreads the columns over the banks.
constexpr int kTileStride = 65;
constexpr int kTileMaxDofs = 16;

// Schedule: after its substep each environment appends itself to one of kSchedClasses queues by cost (its number of
// constraint rows); the next substep concatenates the queues, costliest first, and wavefront w takes the w-th
// environment, so that long environments start early instead of finishing last. The sched field holds the queues of
// both parities (rows [parity * kSchedClasses + class]) and a control row: [0] the parity of the queues the next
// substep reads, [1 + parity * kSchedClasses + class] the queue lengths.
constexpr int kSchedClasses = 4;
__host__ __device__ constexpr int sched_rows() { return 2 * kSchedClasses + 1; }
__device__ __forceinline__ int sched_class(int n_con) {
    return n_con >= 150 ? 0 : n_con >= 60 ? 1 : n_con >= 20 ? 2 : 3;
}

// per-row arrays of the constraint store (each n_constraints floats, after the jacobian)
enum { RW_DIAG = 0, RW_AREF, RW_EFC_D, RW_FL, RW_JAREF, RW_JV, RW_FORCE, RW_ACTIVE, RW_PREV, RW_COUNT };
constexpr int kTmpArrays = 9;

// ------------------------------------------------------------------------------------------- wave primitives
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Make every lane's earlier LDS and global writes visible to the other lanes of this wavefront.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

// exclusive prefix sum over the lanes of the wavefront (all lanes active)
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
    int x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    total = __shfl(x, kWave - 1, kWave);
    return x - v;
}

__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// For every q < NQ: acc[q] + v[q](lane l0) + v[q](lane l0 + 1) + ... + v[q](lane l1 - 1), added in lane order (the
// sequential order of the rows the lanes hold); lane bounds are wave-uniform, the result is the same on every lane.
// The NQ chains are independent, so their adds interleave.
template <int NQ>
__device__ __forceinline__ void lane_seq_sums(float (&acc)[NQ], const float (&v)[NQ], int l0, int l1) {
    if (l0 == 0 && l1 == kWave) {
#pragma unroll
        for (int l = 0; l < kWave; l++)
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] = acc[q] + readlane_f(v[q], l);
    } else {
        for (int l = l0; l < l1; l++)
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] = acc[q] + readlane_f(v[q], l);
    }
}

// acc + f(i0) + f(i0 + 1) + ... + f(i1 - 1), added left to right; up to U terms are loaded before they are added, so
// that the loads of a batch overlap instead of each add waiting for its own load
template <int U = 8, class F>
__device__ __forceinline__ float seq_sum(float acc, int i0, int i1, F f) {
    int i = i0;
    for (; i + U <= i1; i += U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = f(i + u);
#pragma unroll
        for (int u = 0; u < U; u++) acc = acc + v[u];
    }
    for (; i < i1; i++) acc = acc + f(i);
    return acc;
}

// ---------------------------------------------------------------------------------------- private region
template <bool kLds>
__device__ __forceinline__ DState make_local(const DState& st, float* base, const FusedArgs& fa) {
    DState l = st;
    l.B = 1;
    int i = 0;
#define RS_X(name, is_int, kind, width)                                                           \
    if constexpr (is_local_kind(kind)) l.name = (decltype(l.name))(base + fa.loc_off[i]);       \
    i++;
    RS_STATE_FIELDS(RS_X)
#undef RS_X
    return l;
}

// Copy environment b's local rows between the packed local planes (which start at the slab base, the first local
// field) and its private region: local row k is slab element k * n_envs + b. Rows are copied as raw 32-bit words.
template <bool kIn, int kUnroll = 8>
__device__ __forceinline__ void stage_local(const DState& st, uint32_t* priv, int n_local, int b, int lane) {
    uint32_t* slab = (uint32_t*)(void*)st.qpos;
    const size_t B = (size_t)st.B;
    int k = lane;
    for (; k + (kUnroll - 1) * kWave < n_local; k += kUnroll * kWave) {
        uint32_t v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) v[u] = kIn ? slab[(size_t)(k + u * kWave) * B + b] : priv[k + u * kWave];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            if (kIn) priv[k + u * kWave] = v[u];
            else slab[(size_t)(k + u * kWave) * B + b] = v[u];
        }
    }
    for (; k < n_local; k += kWave) {
        if (kIn) priv[k] = slab[(size_t)k * B + b];
        else slab[(size_t)k * B + b] = priv[k];
    }
}

}  // namespace rs
