// Wavefront-per-environment execution for the fused substep.
//
// One 64-lane wavefront owns one environment for the whole substep. The environment's small fields (links, joints,
// dofs, geoms, the mass matrix and solver matrices, SAP buffers, broadphase pairs) are staged from the [row][env]
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
           kind == RS_N_GEOMS || kind == RS_N_ENTITIES || kind == RS_N_DOFS2 || kind == RS_N_SORT ||
           kind == RS_N_BROAD;
}

__host__ __device__ inline int64_t kind_count(const DModel& m, int kind) {
    switch (kind) {
        case RS_N_ONE: return 1;
        case RS_N_LINKS: return m.L;
        case RS_N_JOINTS: return m.J;
        case RS_N_DOFS: return m.D;
        case RS_N_QS: return m.Q;
        case RS_N_GEOMS: return m.G;
        case RS_N_ENTITIES: return m.E;
        case RS_N_EQUALITIES: return m.NE;
        case RS_N_DOFS2: return (int64_t)m.D * m.D;
        case RS_N_GEOMS2: return (int64_t)m.G * m.G;
        case RS_N_PAIRS: return m.P;
        case RS_N_SUPPORT: return m.NSUP;
        case RS_N_CONTACTS: return m.KMAX;
        case RS_N_CONSTRAINTS: return m.CMAX;
        case RS_N_CONSTRAINTS_DOFS: return (int64_t)m.CMAX * m.D;
        case RS_N_BROAD: return m.BROAD;
        case RS_N_SORT: return 2 * (int64_t)m.G;
        default: return 0;
    }
}

// Launch parameters of the fused kernel; offsets are in floats.
struct FusedArgs {
    int epb;           // environments (wavefronts) per workgroup
    int loc_floats;    // private region per environment
    int64_t work_stride;  // floats of the work field per environment
    int off_loc, off_jac, off_rowf, off_tmp, off_pc;
    int loc_off[kMaxStateFields];  // private-region offset of each state field (local kinds only)
};

// per-row arrays of the constraint store (each n_constraints floats, after the jacobian)
enum { RW_DIAG = 0, RW_AREF, RW_EFC_D, RW_FL, RW_JAREF, RW_JV, RW_FORCE, RW_ACTIVE, RW_PREV, RW_COUNT };
constexpr int kTmpArrays = 9;

struct Rows {
    float* jac;  // [c * D + d]
    float* f;    // [k * cap + c], k < RW_COUNT
    float* tmp;  // [k * cap + c], k < kTmpArrays
    int cap;
    __device__ __forceinline__ float& at(int k, int c) const { return f[(size_t)k * cap + c]; }
    __device__ __forceinline__ float& t(int k, int c) const { return tmp[(size_t)k * cap + c]; }
};

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

// copy the local fields of environment b between the planes and the private region
template <bool kIn>
__device__ __forceinline__ void stage_local(const DModel& m, const DState& st, const DState& loc, int b, int lane) {
    const size_t B = (size_t)st.B;
#define RS_X(name, is_int, kind, width)                                                          \
    if constexpr (is_local_kind(kind)) {                                                         \
        const int rows = (int)(kind_count(m, kind) * (width));                                   \
        for (int r = lane; r < rows; r += kWave) {                                               \
            if (kIn) loc.name[r] = st.name[(size_t)r * B + b];                                   \
            else st.name[(size_t)r * B + b] = loc.name[r];                                       \
        }                                                                                        \
    }
    RS_STATE_FIELDS(RS_X)
#undef RS_X
}

}  // namespace rs
