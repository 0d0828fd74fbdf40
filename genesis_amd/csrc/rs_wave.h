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


// Launch parameters of the fused kernel; offsets are in floats.
struct FusedArgs {
    int epb;           // environments (wavefronts) per workgroup
    int n_local;       // local rows staged per environment (the packed local planes)
    int loc_floats;    // private region per environment (n_local rounded up to 16 bytes)
    int tiled;         // row-tile solver class: kTileMaxDofs or kTileWideDofs (the KD its variants are built for), 0: none
    int lds_wave;      // floats of LDS per wavefront: [private region if in LDS][row tile]
    int64_t work_stride;  // floats of the work field per environment
    int off_loc, off_jac, off_rowf, off_tmp, off_pc, off_mc, off_gjk;
    int gjk_words;     // LDS floats per wavefront for the GJK/EPA polytope (0: the polytope stays in the work field)
    int gjk_off;       // offset of that polytope in the wavefront's LDS slice (-1: none)
    int tile_off;      // offset of the row tile (kTileStride x tile_rows(D) floats) in the wavefront's LDS slice
    int sched;         // 1: map wavefronts to environments through the cost-ordered queues
    int loc_off[kMaxStateFields];  // private-region offset of each state field (local kinds only)
};

// Row tile: one chunk of 64 constraint rows staged column by column (column k at k * kTileStride), so that the
// lanes that sum a column over the rows read it from LDS. The stride (68 floats) keeps every row 16-byte aligned, so
// that a full row is read with 16-byte loads, and puts rows q and q + 1 four banks apart, so that up to 16 lanes
// reading different rows at the same column, and 64 lanes writing one row, hit distinct banks.
constexpr int kTileStride = 68;
constexpr int kTileMaxDofs = 16;   // row-tile solver variants for models with up to 16 dofs...
constexpr int kTileWideDofs = 24;  // ... and a second class for 17 to 24 dofs (wider register rows)
constexpr int kGenericMaxThreads = 256;  // lanes per workgroup of the fused kernel for larger models
constexpr int kTileMinRows = 12;   // the ordered sums of the line search use up to 9 rows
__host__ __device__ constexpr int tile_rows(int n_dofs) {
    return n_dofs <= kTileWideDofs && n_dofs + 3 > kTileMinRows ? n_dofs + 3 : kTileMinRows;
}
__host__ __device__ constexpr int tile_class(int n_dofs) {
    return n_dofs <= kTileMaxDofs ? kTileMaxDofs : n_dofs <= kTileWideDofs ? kTileWideDofs : 0;
}

// Schedule: after its substep each environment appends itself to one of kSchedClasses queues by cost (its number of
// constraint rows); the next substep concatenates the queues, costliest first, and wavefront w takes the w-th
// environment, so that long environments start early instead of finishing last. The sched field holds the queues of
// both parities (rows [parity * kSchedClasses + class]) and a control row: [0] the parity of the queues the next
// substep reads, [1 + parity * kSchedClasses + class] the queue lengths.
constexpr int kSchedClasses = 4;
__host__ __device__ constexpr int sched_rows() { return 2 * kSchedClasses + 1; }
constexpr int kSchedHeavyRows = 150;
// gjk_work: the GJK/EPA work of the environment's last narrowphase, lane passes << 16 | single wavefront detections.
// A lane pass (up to 64 detections, one per lane) outweighs any number of rows, so it sets the class when present.
__device__ __forceinline__ int sched_class(int n_con, int gjk_work = 0) {
    const int by_rows = n_con >= kSchedHeavyRows ? 0 : n_con >= 60 ? 1 : n_con >= 20 ? 2 : 3;
    const int passes = gjk_work >> 16, singles = gjk_work & 0xffff;
    if (passes >= 2) return 0;
    if (passes == 1) return min(by_rows, 1);
    if (singles > 0) return min(by_rows, 2);
    return by_rows;
}

// perturbed multi-contact detections of the fused narrowphase: per-pair record and per-detection result (floats)
constexpr int kMcRec = 16;
constexpr int kMcRes = 8;
constexpr int kMcFloats = kWave * kMcRec + 4 * kWave + 4 * kWave * kMcRes;

// per-row arrays of the constraint store (each n_constraints floats, after the jacobian)
enum { RW_DIAG = 0, RW_AREF, RW_EFC_D, RW_FL, RW_JAREF, RW_JV, RW_FORCE, RW_ACTIVE, RW_PREV, RW_COUNT };
constexpr int kTmpArrays = 9;

// ------------------------------------------------------------------------------------------- wave primitives
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// Make every lane's earlier LDS and global writes visible to the other lanes of this wavefront. The fences are at
// wavefront scope: every exchange in the fused kernel stays inside one wavefront (an environment never shares data
// with another wavefront), and the memory model orders a wavefront's own accesses without waiting for them to
// complete, so this orders the code (no compiler motion across it) without stalling on outstanding loads and stores.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// wave_sync for data exchanged through p's memory (LDS or global: the same wavefront-scope ordering)
__device__ __forceinline__ void sync_for(const void*) { wave_sync(); }

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

// acc + row[0] + row[1] + ... + row[63], added in that order; row is 16-byte aligned and read 4 floats per load
__device__ __forceinline__ float row_sum64(float acc, const float* row) {
    float4 v[kWave / 4];
#pragma unroll
    for (int k = 0; k < kWave / 4; k++) v[k] = ((const float4*)row)[k];
#pragma unroll
    for (int k = 0; k < kWave / 4; k++) {
        acc = acc + v[k].x;
        acc = acc + v[k].y;
        acc = acc + v[k].z;
        acc = acc + v[k].w;
    }
    return acc;
}

// acc + a[0] * b[0] + a[1] * b[1] + ... + a[63] * b[63], in that order (each product rounded, then added)
__device__ __forceinline__ float row_dot64(float acc, const float* a, const float* b) {
#pragma unroll
    for (int h = 0; h < 2; h++) {  // two halves of 32 columns bound the registers the loads take
        float4 u[kWave / 8], v[kWave / 8];
#pragma unroll
        for (int k = 0; k < kWave / 8; k++) {
            u[k] = ((const float4*)a)[h * kWave / 8 + k];
            v[k] = ((const float4*)b)[h * kWave / 8 + k];
        }
#pragma unroll
        for (int k = 0; k < kWave / 8; k++) {
            acc = acc + u[k].x * v[k].x;
            acc = acc + u[k].y * v[k].y;
            acc = acc + u[k].z * v[k].z;
            acc = acc + u[k].w * v[k].w;
        }
    }
    return acc;
}

// Ordered sums through the row tile. Lane l writes its row values v[q] to column l of tile row q; lane q (q < NQ) then
// adds row q's columns l0, l0 + 1, ..., l1 - 1 in that order to its own accumulator acc: the additions of a sequential
// loop over the rows, with the NQ chains running side by side in NQ lanes instead of as NQ x 64 broadcasts. The row
// stride kTileStride keeps both the lane-wise writes and the row-wise reads free of bank conflicts. Bounds are
// wave-uniform; the accumulators stay in lanes 0..NQ-1 (read them back with readlane_f).
template <int NQ>
__device__ __forceinline__ void tile_seq_sums(float& acc, const float (&v)[NQ], int l0, int l1, float* tile, int lane) {
#pragma unroll
    for (int q = 0; q < NQ; q++) tile[q * kTileStride + lane] = v[q];
    sync_for(tile);
    if (lane < NQ) {
        const float* row = tile + lane * kTileStride;
        if (l0 == 0 && l1 == kWave) {
            acc = row_sum64(acc, row);
        } else {
            acc = seq_sum(acc, l0, l1, [&](int l) { return row[l]; });
        }
    }
    sync_for(tile);
}

// ---------------------------------------------------------------------------------------- private region
// the local view of the environment whose private region starts at base
__device__ __forceinline__ LState make_local(float* base, const FusedArgs& fa) {
    return LState{base, (const RS_CONST int*)fa.loc_off};
}

// Copy environment b's local rows between the slab and its private region, as raw 32-bit words. With env-major
// records (st.LS > 0) the rows are one contiguous, 128-byte aligned record (copied 16 bytes per lane and
// instruction); with planes, local row k is slab element k * n_envs + b (the packed local planes start at the slab
// base, the first local field).
template <bool kIn, int kUnroll = 8>
__device__ __forceinline__ void stage_local(const DState& st, uint32_t* priv, int n_local, int b, int lane) {
    uint32_t* slab = (uint32_t*)(void*)st.qpos;
    if (st.LS > 0) {
        uint32_t* rec = slab + (size_t)b * st.LS;
        const int n4 = n_local >> 2;
        int k = lane;
        for (; k + (kUnroll - 1) * kWave < n4; k += kUnroll * kWave) {
            uint4 v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) v[u] = kIn ? ((const uint4*)rec)[k + u * kWave] : ((const uint4*)priv)[k + u * kWave];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                if (kIn) ((uint4*)priv)[k + u * kWave] = v[u];
                else ((uint4*)rec)[k + u * kWave] = v[u];
            }
        }
        for (; k < n4; k += kWave) {
            if (kIn) ((uint4*)priv)[k] = ((const uint4*)rec)[k];
            else ((uint4*)rec)[k] = ((const uint4*)priv)[k];
        }
        for (int t = 4 * n4 + lane; t < n_local; t += kWave) {
            if (kIn) priv[t] = rec[t];
            else rec[t] = priv[t];
        }
        return;
    }
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
