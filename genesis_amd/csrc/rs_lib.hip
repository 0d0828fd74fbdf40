// librigidstep: C ABI (include/rigidstep.h) over the gfx950 rigid-step kernels.
//
// Plain path: one kernel per stage of RigidSolver.substep (rigid_solver.py:912-980), one lane per environment, every
// intermediate kept in the [row][env] planes of the state slab (the comparison point and drop-in).
// Fused path: the whole substep in a single kernel per launch, one wavefront per environment (rs_fused.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rs_collision.h"
#include "rs_constraint.h"
#include "rs_dyn.h"
#include "rs_fused.h"
#include "rs_launch.h"

using namespace rs;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_OK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(RS_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------------------------ directories
namespace {
struct Info {
    const char* name;
    int is_int, count, width;
};
const Info k_model_arrays[] = {
#define RS_X(name, is_int, kind, width) {#name, is_int, kind, width},
    RS_MODEL_ARRAYS(RS_X)
#undef RS_X
};
const Info k_state_fields[] = {
#define RS_X(name, is_int, kind, width) {#name, is_int, kind, width},
    RS_STATE_FIELDS(RS_X)
#undef RS_X
};
constexpr int k_n_model_arrays = sizeof(k_model_arrays) / sizeof(Info);
constexpr int k_n_state_fields = sizeof(k_state_fields) / sizeof(Info);
}  // namespace

struct rs_model {
    rs_model_header h;
    void* blob = nullptr;
    size_t blob_bytes = 0;
    DModel dm;
    int device = 0;
    FusedArgs fa;       // fused-kernel layout and launch shape
    int use_lds = 0;    // private regions in LDS
    int schedule = 1;   // cost-ordered environment queues
    int occupancy = 0;  // 0: compiler's register budget; 3 or 4: at least that many wavefronts per SIMD
    int layout = RS_LAYOUT_ENV_RECORDS;  // state layout of the per-environment (local-kind) fields
    int max_lds = 0;    // device limit of dynamic LDS per workgroup (bytes)
    int n_cu = 0;       // compute units of the device
    bool has_mesh = false;  // convex mesh geoms: the GJK/EPA polytope may be kept in LDS (gjk_lds)
};

// Shared launches of several (model, slab) batches: a ring of pinned host + device segment tables, each reused only
// after the copy that last read it has completed.
struct rs_batch {
    static constexpr int kRing = 4;
    int max_segments = 0;
    int slot = 0;
    FusedSegment* host[kRing] = {};
    FusedSegment* dev[kRing] = {};
    hipEvent_t done[kRing] = {};
    bool pending[kRing] = {};
};

struct rs_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int64_t work_floats(const rs_model_header& h);
static int64_t field_rows(const rs_model_header& h, int i);

static int64_t count_of(const rs_model_header& h, int kind) {
    switch (kind) {
        case RS_N_ONE: return 1;
        case RS_N_LINKS: return h.n_links;
        case RS_N_JOINTS: return h.n_joints;
        case RS_N_DOFS: return h.n_dofs;
        case RS_N_QS: return h.n_qs;
        case RS_N_GEOMS: return h.n_geoms;
        case RS_N_ENTITIES: return h.n_entities;
        case RS_N_EQUALITIES: return h.n_equalities;
        case RS_N_DOFS2: return (int64_t)h.n_dofs * h.n_dofs;
        case RS_N_GEOMS2: return (int64_t)h.n_geoms * h.n_geoms;
        case RS_N_PAIRS: return h.n_pairs;
        case RS_N_SUPPORT: return h.n_support;
        case RS_N_VERTS: return h.n_verts;
        case RS_N_CONTACTS: return h.max_contact_pairs;
        case RS_N_CONSTRAINTS: return h.len_constraints;
        case RS_N_CONSTRAINTS_DOFS: return (int64_t)h.len_constraints * h.n_dofs;
        case RS_N_BROAD: return h.max_collision_pairs_broad;
        case RS_N_SORT: return 2 * (int64_t)h.n_geoms;
        case RS_N_WORK: return work_floats(h);
        case RS_N_SCHED: return sched_rows();
        case RS_N_DOFS_PLANE: return h.n_dofs;
        case RS_N_ONE_PLANE: return 1;
        default: return 0;
    }
}

// rows per environment of state field i (at least 1, so that every plane is addressable even for empty models)
static int64_t field_rows(const rs_model_header& h, int i) {
    int64_t rows = count_of(h, k_state_fields[i].count) * k_state_fields[i].width;
    return rows < 1 ? 1 : rows;
}

// Private-region offsets of the local fields and the env-major work layout of the fused kernel (floats, each
// sub-array 16-byte aligned, each environment's slice 256-byte aligned).
static void fused_layout(const rs_model_header& h, FusedArgs* fa) {
    auto up4 = [](int64_t x) { return (x + 3) / 4 * 4; };
    // local fields come first in the directory; the private region mirrors their packed planes row for row
    int64_t off = 0;
    for (int i = 0; i < k_n_state_fields && i < kMaxStateFields; i++) {
        fa->loc_off[i] = -1;
        if (!is_local_kind(k_state_fields[i].count)) continue;
        fa->loc_off[i] = (int)off;
        off += field_rows(h, i);
    }
    fa->n_local = (int)off;
    off = up4(off);
    fa->loc_floats = (int)off;
    fa->tiled = tile_class(h.n_dofs);
    fa->tile_off = 0;  // set per launch by set_lds_slice
    fa->gjk_words = 0;  // set at model creation (models with mesh geoms)
    fa->gjk_off = -1;
    fa->lds_wave = (int)up4((int64_t)kTileStride * tile_rows(h.n_dofs));
    fa->sched = 0;
    const int64_t C = h.len_constraints > 0 ? h.len_constraints : 1, D = h.n_dofs;
    fa->off_loc = 0;
    fa->off_jac = (int)up4(off);
    fa->off_rowf = (int)(fa->off_jac + up4(C * D));
    fa->off_tmp = (int)(fa->off_rowf + RW_COUNT * up4(C));
    fa->off_pc = (int)(fa->off_tmp + kTmpArrays * up4(C));
    fa->off_mc = (int)(fa->off_pc + up4((int64_t)kWave * (h.n_contacts_per_pair > 0 ? h.n_contacts_per_pair : 1) * 8));
    // GJK/EPA polytopes: one at the reference's capacities (the per-stage kernels use it), one compact one per lane
    fa->off_gjk = (int)up4(fa->off_mc + kMcFloats);
    int64_t total = fa->off_gjk + (int64_t)kGjkWorkWords;
    fa->work_stride = (total + 63) / 64 * 64;
}

// LDS slice of one wavefront: the private region (when it is in LDS) followed by the row tile
static int tile_floats(const rs_model_header& h) { return (kTileStride * tile_rows(h.n_dofs) + 3) / 4 * 4; }
// slice layout: [private region (LDS mode)][row tile][GJK/EPA polytope (models with mesh geoms)]
static void set_lds_slice(const rs_model_header& h, FusedArgs* fa, bool use_lds, int slice_floats) {
    fa->tile_off = use_lds ? fa->loc_floats : 0;
    fa->gjk_off = fa->gjk_words > 0 ? fa->tile_off + tile_floats(h) : -1;
    fa->lds_wave = slice_floats;
}
static int lds_slice_floats(const rs_model_header& h, const FusedArgs& fa, bool use_lds) {
    return (use_lds ? fa.loc_floats : 0) + tile_floats(h) + fa.gjk_words;
}

static int64_t work_floats(const rs_model_header& h) {
    FusedArgs fa;
    fused_layout(h, &fa);
    return fa.work_stride;
}

// ------------------------------------------------------------------------------------------------ state views
// Local-kind fields come first in the directory. RS_LAYOUT_PLANES: they are [rows][n_envs] planes packed back to
// back. RS_LAYOUT_ENV_RECORDS: each environment's local rows form one record of record_floats() elements (the rows
// of the local fields back to back, padded to 128 bytes) and the records follow each other; offsets[i] of a local
// field is then its byte offset inside a record. Every other field is a [rows][n_envs] plane starting 256-byte
// aligned after them.
static int64_t local_rows(const rs_model_header& h) {
    int64_t n = 0;
    for (int i = 0; i < k_n_state_fields; i++)
        if (is_local_kind(k_state_fields[i].count)) n += field_rows(h, i);
    return n;
}
static int64_t record_floats(const rs_model* m) {
    return m->layout == RS_LAYOUT_ENV_RECORDS ? (local_rows(m->h) + 31) / 32 * 32 : 0;
}

static void layout(const rs_model* m, int n_envs, int64_t* offs, int64_t* total) {
    const int64_t rec = record_floats(m);
    int64_t off = 0, in_rec = 0;
    for (int i = 0; i < k_n_state_fields; i++) {
        const int64_t rows = field_rows(m->h, i);
        const bool local = is_local_kind(k_state_fields[i].count);
        if (local && rec > 0) {
            offs[i] = in_rec * 4;
            in_rec += rows;
            continue;
        }
        if (!local) {
            if (off == 0 && rec > 0) off = rec * n_envs * 4;
            off = (off + 255) / 256 * 256;
        }
        offs[i] = off;
        off += rows * n_envs * 4;
    }
    *total = (off + 255) / 256 * 256;
}

static DState make_state(const rs_model* m, void* base, int n_envs) {
    int64_t offs[k_n_state_fields];
    int64_t total;
    layout(m, n_envs, offs, &total);
    DState s;
    s.B = n_envs;
    s.LS = record_floats(m);
    int i = 0;
    char* p = (char*)base;
#define RS_X(name, is_int, kind, width) s.name = (decltype(s.name))(p + offs[i++]);
    RS_STATE_FIELDS(RS_X)
#undef RS_X
    return s;
}

// ------------------------------------------------------------------------------------------------- kernels
static constexpr int kBlock = 64;

__global__ void __launch_bounds__(kBlock) k_init(DModel m, DState st) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= st.B) return;
    for (int q = 0; q < m.Q; q++) SF(qpos, q) = m.qpos0[q];
    for (int d = 0; d < m.D; d++) SF(ctrl_mode, d) = CTRL_FORCE;
    SF(first_time, 0) = 1;
    for (int l = 0; l < m.L; l++) {
        ST3(links_pos, l, mvec3(m.link_pos, l));
        ST4(links_quat, l, mquat(m.link_quat, l));
    }
    dev_kinematics(m, st, b, true);
}

__global__ void __launch_bounds__(kBlock) k_fk_masked(DModel m, DState st, const int32_t* mask) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= st.B) return;
    if (mask && !mask[b]) return;
    dev_kinematics(m, st, b, false);
}

__global__ void __launch_bounds__(kBlock) k_reset_caches(DModel m, DState st, const int32_t* mask) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= st.B) return;
    if (mask && !mask[b]) return;
    for (int d = 0; d < m.D; d++) SF(qacc_ws, d) = 0.f;
    SF(is_warmstart, 0) = 0;
    for (int p = 0; p < 3 * m.P; p++) SF(cache_normal, p) = 0.f;
    SF(first_time, 0) = 1;
    SF(errbits, 0) = 0;
}

// one kernel per stage (distinct symbols, so kernel traces attribute time per stage)
template <int stage>
__global__ void __launch_bounds__(kBlock) k_stage(DModel m, DState st) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= st.B) return;
    switch (stage) {
        case RS_STAGE_FK: dev_kinematics(m, st, b, false); break;
        case RS_STAGE_CRBA: dev_mass_matrix(m, st, b); break;
        case RS_STAGE_FACTOR: dev_factor_mass(m, st, b); break;
        case RS_STAGE_RNE:
            dev_torque_passive(m, st, b);
            dev_update_acc(m, st, b, false);
            dev_rne_and_qacc(m, st, b);
            break;
        case RS_STAGE_BROADPHASE: dev_broadphase(m, st, b); break;
        case RS_STAGE_NARROWPHASE: dev_narrowphase(m, st, b); break;
        case RS_STAGE_CONSTRAINTS:
            dev_equality_constraints(m, st, b);
            dev_inequality_constraints(m, st, b);
            break;
        case RS_STAGE_SOLVE: dev_solve(m, st, b); break;
        case RS_STAGE_INTEGRATE:
            dev_update_acc(m, st, b, true);
            if (m.integrator != RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) dev_implicit_damping(m, st, b);
            dev_integrate(m, st, b);
            dev_kinematics(m, st, b, false);
            break;
        default: break;
    }
}

// one lane per (geom, env)
__global__ void __launch_bounds__(256) k_aabb(DModel m, DState st) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)m.G * st.B) return;
    int g = (int)(t / st.B), b = (int)(t % st.B);
    dev_aabb(m, st, b, g);
}

__global__ void k_errno_or(DState st, int32_t* out) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    int v = b < st.B ? FV(st, errbits, b)[0] : 0;
    if (v) atomicOr(out, v);
}

// ------------------------------------------------------------------------------- single detections (rs_pair_contact)
__device__ __forceinline__ bool pair_item(const DModel& m, int n, int i, const int32_t* ga, const int32_t* gb,
                                          const float* pa, const float* pb, GjkPair& p) {
    if (i >= n) return false;
    const int g1 = ga[i], g2 = gb[i];
    if (g1 < 0 || g1 >= m.G || g2 < 0 || g2 >= m.G) return false;
    const int ta = m.geom_type[g1], tb = m.geom_type[g2];
    if (ta == GT_PLANE || tb == GT_PLANE || ta == GT_TERRAIN || tb == GT_TERRAIN) return false;
    const float* a = pa + 7 * (size_t)i;
    const float* c = pb + 7 * (size_t)i;
    p = make_gjk_pair(m, g1, g2, mk3(a[0], a[1], a[2]), f4q{a[3], a[4], a[5], a[6]}, mk3(c[0], c[1], c[2]),
                      f4q{c[3], c[4], c[5], c[6]});
    return true;
}

__device__ __forceinline__ void pair_out(int i, bool valid, bool col, f3 nrm, f3 pos, float pen, float* out,
                                         int32_t* hit) {
    float* o = out + 7 * (size_t)i;
    o[0] = nrm.x; o[1] = nrm.y; o[2] = nrm.z;
    o[3] = pos.x; o[4] = pos.y; o[5] = pos.z;
    o[6] = pen;
    hit[i] = valid ? (col ? 1 : 0) : -1;
}

// per-item floats of rs_pair_contact's work buffer: a polytope at the reference's capacities plus a compact one
static constexpr int kPairWorkWords = kGjkMemSlotWords + kGjkLdsWords;

// one wavefront per item: the fused kernel's GJK/EPA for a single detection (polytope in LDS)
__global__ void __launch_bounds__(kWave) k_pair_gjk_wave(DModel m, int n, const int32_t* ga, const int32_t* gb,
                                                         const float* pa, const float* pb, float* work, float* out,
                                                         int32_t* hit) {
    __shared__ __align__(16) float lds[kGjkLdsWords];
    const int i = blockIdx.x;
    GjkPair p;
    const bool valid = pair_item(m, n, i, ga, gb, pa, pb, p);
    if (i >= n) return;
    f3 nrm = zero3(), pos = zero3();
    float pen = 0.f;
    const bool col =
        valid ? gjk_contact_wave(m, p, lane_id(), lds, work + (size_t)i * kPairWorkWords, nrm, pos, pen) : false;
    if (lane_id() == 0) pair_out(i, valid, col, nrm, pos, pen, out, hit);
}

// one lane per item: the fused kernel's lane-parallel GJK/EPA (compact polytope per lane, the block's full polytope for
// detections that outgrow it; the block of 64 items owns work[64 * blockIdx.x * kPairWorkWords ...]) or MPR without
// a warm-start direction
__global__ void __launch_bounds__(kBlock) k_pair_lane(DModel m, int n, const int32_t* ga, const int32_t* gb,
                                                      const float* pa, const float* pb, int method, float* work,
                                                      float* out, int32_t* hit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    GjkPair p;
    const bool valid = pair_item(m, n, i, ga, gb, pa, pb, p);
    if (i >= n) return;
    f3 nrm = zero3(), pos = zero3();
    float pen = 0.f;
    bool col = false;
    if (method == RS_PAIR_GJK) {
        gjk_contact_lanes(m, p, valid, lane_id(), work + (size_t)blockIdx.x * kBlock * kPairWorkWords, col, nrm, pos,
                          pen);
    } else if (valid) {
        mpr_contact(m, p.ga, p.gb, zero3(), p.pa, p.qa, p.pb, p.qb, col, nrm, pen, pos);
    }
    pair_out(i, valid, col, nrm, pos, pen, out, hit);
}

// ---------------------------------------------------------------------- error-word reduction without allocation
// rs_errno_reduce_noalloc keeps its partial results in `scratch`, a file-scope device array (each GPU holds its own
// copy until the process ends): one partial OR per block of kErrnoBlock environments in words [0, n_blocks), the
// final OR in word kErrnoMaxPartials. Every call rewrites every word it reads, so nothing has to start at zero; the
// array keeps the last call's contents, and calls that overlap on one GPU race.
constexpr int kErrnoBlock = 256;
constexpr int kErrnoMaxPartials = 4096;  // at most kErrnoMaxPartials * kErrnoBlock = 1,048,576 environments per call
static __device__ __align__(256) int32_t scratch[kErrnoMaxPartials + 64];

__device__ __forceinline__ int32_t block_or(int32_t v, int32_t* w) {
    for (int off = 1; off < kWave; off <<= 1) v |= __shfl_xor(v, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) w[threadIdx.x / kWave] = v;
    __syncthreads();
    int32_t r = 0;
    for (int i = 0; i < kErrnoBlock / kWave; i++) r |= w[i];
    return r;
}

// OR of block blockIdx.x's environments' error words into partials[blockIdx.x]
__global__ void __launch_bounds__(kErrnoBlock) k_errno_partial(DState st, int32_t* partials) {
    __shared__ int32_t w[kErrnoBlock / kWave];
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t r = block_or(b < st.B ? FV(st, errbits, b)[0] : 0, w);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// OR of n partial words into *out (one block)
__global__ void __launch_bounds__(kErrnoBlock) k_errno_final(const int32_t* partials, int n, int32_t* out) {
    __shared__ int32_t w[kErrnoBlock / kWave];
    int32_t v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v |= partials[i];
    const int32_t r = block_or(v, w);
    if (threadIdx.x == 0) *out = r;
}

// ------------------------------------------------------------------------------------------------ launch helpers
static inline dim3 env_grid(int n) { return dim3((n + kBlock - 1) / kBlock); }

// Environments per workgroup of the narrowphase kernel (one lane per environment): the largest power of two up to 64
// that still gives every compute unit four workgroups, so that a batch spreads over the whole device and each
// wavefront waits on fewer environments' divergent detections (4096 environments on 256 CUs: 4 per workgroup). The
// other per-stage kernels keep 64 environments per workgroup, whose loads of [row][env] planes coalesce.
static inline int plain_block(const rs_model* m, int n_envs) {
    const int n_cu = m->n_cu > 0 ? m->n_cu : 1;
    int tpb = kBlock;
    while (tpb > 1 && (n_envs + tpb - 1) / tpb < 4 * n_cu) tpb >>= 1;
    return tpb;
}

static int launch_stage(const rs_model* m, const DState& st, int stage, hipStream_t s) {
    const DModel& dm = m->dm;
    if (stage == RS_STAGE_AABB) {
        int64_t n = (int64_t)dm.G * st.B;
        if (n > 0) hipLaunchKernelGGL(k_aabb, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dm, st);
    } else if ((stage == RS_STAGE_BROADPHASE || stage == RS_STAGE_NARROWPHASE) && dm.P == 0) {
        // no candidate pairs: detection stops after the AABBs (collider.py:690-691)
    } else if ((stage == RS_STAGE_CONSTRAINTS || stage == RS_STAGE_SOLVE) && dm.disable_constraint) {
    } else {
        const int tpb = stage == RS_STAGE_NARROWPHASE ? plain_block(m, st.B) : kBlock;
        dim3 g((unsigned)((st.B + tpb - 1) / tpb)), blk((unsigned)tpb);
        switch (stage) {
            case RS_STAGE_FK: hipLaunchKernelGGL(k_stage<RS_STAGE_FK>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_CRBA: hipLaunchKernelGGL(k_stage<RS_STAGE_CRBA>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_FACTOR: hipLaunchKernelGGL(k_stage<RS_STAGE_FACTOR>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_RNE: hipLaunchKernelGGL(k_stage<RS_STAGE_RNE>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_BROADPHASE: hipLaunchKernelGGL(k_stage<RS_STAGE_BROADPHASE>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_NARROWPHASE: hipLaunchKernelGGL(k_stage<RS_STAGE_NARROWPHASE>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_CONSTRAINTS: hipLaunchKernelGGL(k_stage<RS_STAGE_CONSTRAINTS>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_SOLVE: hipLaunchKernelGGL(k_stage<RS_STAGE_SOLVE>, g, blk, 0, s, dm, st); break;
            case RS_STAGE_INTEGRATE: hipLaunchKernelGGL(k_stage<RS_STAGE_INTEGRATE>, g, blk, 0, s, dm, st); break;
            default: break;
        }
    }
    HIP_OK(hipGetLastError());
    return RS_OK;
}

static int fused_substep(const rs_model* m, const DState& st, hipStream_t s) {
    const int epb = m->fa.epb;
    FusedArgs fa = m->fa;
    fa.sched = m->schedule && st.B >= sched_rows();
    if (fa.sched) hipLaunchKernelGGL(k_sched_flip, dim3(1), dim3(1), 0, s, st);
    const dim3 grid((unsigned)((st.B + epb - 1) / epb)), block((unsigned)(kWave * epb));
    set_lds_slice(m->h, &fa, m->use_lds, lds_slice_floats(m->h, fa, m->use_lds));
    const size_t lds = (size_t)epb * fa.lds_wave * 4;
    // one kernel variant per (private region, occupancy) for models with a row tile; models without one run the
    // generic-solver variants at the compiler's register budget (each class is instantiated in its own translation
    // unit: rs_fused_t16.hip, rs_fused_t24.hip, rs_fused_gen.hip)
    const bool lds_on = m->use_lds != 0;
    if (fa.tiled == kTileMaxDofs) launch_fused_tile16(lds_on, m->occupancy, grid, block, lds, s, m->dm, st, fa);
    else if (fa.tiled == kTileWideDofs) launch_fused_tile24(lds_on, m->occupancy, grid, block, lds, s, m->dm, st, fa);
    else launch_fused_generic(lds_on, grid, block, lds, s, m->dm, st, fa);
    HIP_OK(hipGetLastError());
    return RS_OK;
}

static int check_fused_config(const rs_model* m, const FusedArgs& fa, int epb, int use_lds) {
    if (epb < 1 || epb > 8) return fail(RS_E_INVALID, "fused config: envs_per_block must be in 1..8");
    if (!m->fa.tiled && epb * kWave > kGenericMaxThreads)
        return fail(RS_E_UNSUPPORTED, "fused config: models above 24 dofs run at most 4 environments per workgroup");
    if ((int64_t)epb * lds_slice_floats(m->h, fa, use_lds) * 4 > m->max_lds)
        return fail(RS_E_UNSUPPORTED, use_lds ? "fused config: private regions do not fit in LDS at this envs_per_block"
                                              : "fused config: row tiles do not fit in LDS at this envs_per_block");
    return RS_OK;
}

static int plain_substep(const rs_model* m, const DState& st, hipStream_t s) {
    const int order[] = {RS_STAGE_CRBA, RS_STAGE_FACTOR, RS_STAGE_RNE};
    for (int stg : order)
        if (int r = launch_stage(m, st, stg, s)) return r;
    if (m->dm.enable_collision) {
        for (int stg : {RS_STAGE_AABB, RS_STAGE_BROADPHASE, RS_STAGE_NARROWPHASE})
            if (int r = launch_stage(m, st, stg, s)) return r;
    }
    if (!m->dm.disable_constraint) {
        for (int stg : {RS_STAGE_CONSTRAINTS, RS_STAGE_SOLVE})
            if (int r = launch_stage(m, st, stg, s)) return r;
    }
    return launch_stage(m, st, RS_STAGE_INTEGRATE, s);
}

// ------------------------------------------------------------------------------------------------------- ABI
extern "C" {

const char* rs_last_error(void) { return g_err.c_str(); }
int rs_abi_version(void) { return RS_ABI_VERSION; }

int rs_model_array_count(void) { return k_n_model_arrays; }
int rs_model_array_info(int i, rs_array_info* out) {
    if (i < 0 || i >= k_n_model_arrays || !out) return fail(RS_E_INVALID, "rs_model_array_info: bad index");
    out->name = k_model_arrays[i].name;
    out->is_int = k_model_arrays[i].is_int;
    out->count = k_model_arrays[i].count;
    out->width = k_model_arrays[i].width;
    return RS_OK;
}
int rs_state_field_count(void) { return k_n_state_fields; }
int rs_state_field_info(int i, rs_array_info* out) {
    if (i < 0 || i >= k_n_state_fields || !out) return fail(RS_E_INVALID, "rs_state_field_info: bad index");
    out->name = k_state_fields[i].name;
    out->is_int = k_state_fields[i].is_int;
    out->count = k_state_fields[i].count;
    out->width = k_state_fields[i].width;
    return RS_OK;
}

// whether the host blob holds a convex mesh geom (geom_type == GT_MESH and geom_is_convex)
static bool has_convex_mesh(const rs_model_header& h, const char* blob) {
    const int32_t *type = nullptr, *convex = nullptr;
    for (int i = 0; i < k_n_model_arrays; i++) {
        if (!strcmp(k_model_arrays[i].name, "geom_type")) type = (const int32_t*)(blob + h.array_offset[i]);
        if (!strcmp(k_model_arrays[i].name, "geom_is_convex")) convex = (const int32_t*)(blob + h.array_offset[i]);
    }
    if (!type || !convex) return false;
    for (int g = 0; g < h.n_geoms; g++)
        if (type[g] == GT_MESH && convex[g]) return true;
    return false;
}

int rs_model_create(const rs_model_header* h, const void* blob, size_t blob_bytes, rs_model** out) {
    if (!h || !blob || !out) return fail(RS_E_INVALID, "rs_model_create: null argument");
    if (h->abi_version != RS_ABI_VERSION) return fail(RS_E_VERSION, "rs_model_create: ABI version mismatch");
    if (h->float_bytes != 4) return fail(RS_E_INVALID, "rs_model_create: device blob must hold float32");
    if (h->n_arrays != k_n_model_arrays) return fail(RS_E_VERSION, "rs_model_create: array directory mismatch");
    if (h->n_dofs < 0 || h->n_links < 1 || h->n_geoms < 0) return fail(RS_E_INVALID, "rs_model_create: bad counts");
    for (int i = 0; i < k_n_model_arrays; i++) {
        int64_t n = count_of(*h, k_model_arrays[i].count) * k_model_arrays[i].width * 4;
        if (h->array_offset[i] < 0 || h->array_offset[i] + n > (int64_t)blob_bytes)
            return fail(RS_E_INVALID, std::string("rs_model_create: array out of blob: ") + k_model_arrays[i].name);
    }
    rs_model* m = new rs_model();
    m->h = *h;
    HIP_OK(hipGetDevice(&m->device));
    m->blob_bytes = blob_bytes;
    hipError_t e = hipMalloc(&m->blob, blob_bytes);
    if (e != hipSuccess) {
        delete m;
        return fail(RS_E_DEVICE, std::string("hipMalloc model: ") + hipGetErrorString(e));
    }
    e = hipMemcpy(m->blob, blob, blob_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(m->blob);
        delete m;
        return fail(RS_E_DEVICE, std::string("hipMemcpy model: ") + hipGetErrorString(e));
    }
    DModel& d = m->dm;
    d.L = h->n_links; d.J = h->n_joints; d.D = h->n_dofs; d.Q = h->n_qs; d.G = h->n_geoms; d.E = h->n_entities;
    d.NE = h->n_equalities; d.P = h->n_pairs; d.NSUP = h->n_support; d.KMAX = h->max_contact_pairs;
    d.CMAX = h->len_constraints; d.BROAD = h->max_collision_pairs_broad; d.NCPP = h->n_contacts_per_pair;
    d.dt = h->dt; d.gx = h->gravity[0]; d.gy = h->gravity[1]; d.gz = h->gravity[2];
    d.tolerance = h->tolerance; d.ls_tolerance = h->ls_tolerance; d.meaninertia = h->meaninertia;
    d.mc_perturbation = h->mc_perturbation; d.mc_tolerance = h->mc_tolerance;
    d.mpr_to_gjk_overlap_ratio = h->mpr_to_gjk_overlap_ratio;
    d.eps = 1.1920928955078125e-07f;
    d.iterations = h->iterations; d.ls_iterations = h->ls_iterations;
    d.enable_collision = h->enable_collision; d.enable_joint_limit = h->enable_joint_limit;
    d.disable_constraint = h->disable_constraint; d.enable_multi_contact = h->enable_multi_contact;
    d.box_box_detection = h->box_box_detection; d.integrator = h->integrator; d.solver_type = h->solver_type;
    int i = 0;
    const char* base = (const char*)m->blob;
#define RS_X(name, is_int, kind, width) d.name = (decltype(d.name))(base + h->array_offset[i++]);
    RS_MODEL_ARRAYS(RS_X)
#undef RS_X
    if (h->solver_type != RS_SOLVER_CG && h->solver_type != RS_SOLVER_NEWTON) {
        rs_model_destroy(m);
        return fail(RS_E_INVALID, "rs_model_create: solver_type must be RS_SOLVER_CG or RS_SOLVER_NEWTON");
    }
    if (h->integrator != RS_INTEGRATOR_EULER && h->integrator != RS_INTEGRATOR_IMPLICITFAST &&
        h->integrator != RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) {
        rs_model_destroy(m);
        return fail(RS_E_INVALID, "rs_model_create: integrator must be one of RS_INTEGRATOR_*");
    }
    static_assert(k_n_state_fields <= kMaxStateFields, "FusedArgs::loc_off too small");
    fused_layout(m->h, &m->fa);
    d.work_stride = m->fa.work_stride;
    d.off_gjk = m->fa.off_gjk;
    // models with convex mesh geoms meet the GJK/EPA hand-over often (deep overlaps of hulls): their wavefronts may
    // keep a small polytope in LDS (gjk_lds); by default it stays in the work field, which leaves LDS for more
    // wavefronts per compute unit (Franka: 11.5 instead of 17.9 KB per environment)
    m->has_mesh = has_convex_mesh(*h, (const char*)blob);
    m->fa.gjk_words = 0;
    int max_lds = 0;
    if (hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, m->device) != hipSuccess) max_lds = 0;
    m->max_lds = max_lds < 65536 ? max_lds : 65536;
    if (hipDeviceGetAttribute(&m->n_cu, hipDeviceAttributeMultiprocessorCount, m->device) != hipSuccess) m->n_cu = 0;
    // default launch shape: two environments per workgroup, private regions in LDS when they fit
    m->fa.epb = 2;
    m->use_lds = check_fused_config(m, m->fa, 2, 1) == RS_OK;
    g_err.clear();
    *out = m;
    return RS_OK;
}

int rs_model_update(rs_model* m, const rs_model_header* h, const void* blob, size_t blob_bytes, void* stream) {
    if (!m || !h || !blob) return fail(RS_E_INVALID, "rs_model_update: null argument");
    if (blob_bytes != m->blob_bytes || std::memcmp(h, &m->h, sizeof(rs_model_header)) != 0)
        return fail(RS_E_INVALID, "rs_model_update: the header or blob size differs from the model's (only array "
                                  "values may change)");
    hipStream_t s = (hipStream_t)stream;
    HIP_OK(hipMemcpyAsync(m->blob, blob, blob_bytes, hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    return RS_OK;
}

int rs_fused_config_get(const rs_model* m, rs_fused_config* out) {
    if (!m || !out) return fail(RS_E_INVALID, "rs_fused_config_get: null argument");
    out->envs_per_block = m->fa.epb;
    out->use_lds = m->use_lds;
    out->lds_bytes_per_env = lds_slice_floats(m->h, m->fa, m->use_lds) * 4;
    out->schedule = m->schedule;
    out->occupancy = m->occupancy;
    out->gjk_lds = m->fa.gjk_words > 0;
    return RS_OK;
}

int rs_fused_config_set(rs_model* m, const rs_fused_config* c) {
    if (!m || !c) return fail(RS_E_INVALID, "rs_fused_config_set: null argument");
    if (c->gjk_lds && !m->has_mesh)
        return fail(RS_E_UNSUPPORTED, "fused config: gjk_lds applies to models with convex mesh geoms only");
    FusedArgs fa = m->fa;
    fa.gjk_words = c->gjk_lds ? kGjkLdsWords : 0;
    if (int r = check_fused_config(m, fa, c->envs_per_block, c->use_lds != 0)) return r;
    if (!m->fa.tiled && c->occupancy != 0)
        return fail(RS_E_UNSUPPORTED, "fused config: occupancy variants exist for models with a row tile only");
    if (m->fa.tiled == kTileWideDofs && c->occupancy == 4)
        return fail(RS_E_UNSUPPORTED, "fused config: models with 17 to 24 dofs have occupancy variants 0 and 3 only");
    m->fa.epb = c->envs_per_block;
    m->use_lds = c->use_lds != 0;
    m->schedule = c->schedule != 0;
    m->occupancy = (c->occupancy == 3 || c->occupancy == 4) ? c->occupancy : 0;
    m->fa.gjk_words = fa.gjk_words;
    return RS_OK;
}

int rs_state_field_rows(const rs_model* m, int i, int64_t* rows) {
    if (!m || !rows || i < 0 || i >= k_n_state_fields) return fail(RS_E_INVALID, "rs_state_field_rows: bad argument");
    *rows = field_rows(m->h, i);
    return RS_OK;
}

void rs_model_destroy(rs_model* m) {
    if (!m) return;
    if (m->blob) (void)hipFree(m->blob);
    delete m;
}

int rs_state_layout(const rs_model* m, int n_envs, int64_t* offsets, int64_t* total_bytes) {
    if (!m || n_envs < 1 || !total_bytes) return fail(RS_E_INVALID, "rs_state_layout: bad argument");
    std::vector<int64_t> offs(k_n_state_fields);
    layout(m, n_envs, offs.data(), total_bytes);
    if (offsets) std::memcpy(offsets, offs.data(), sizeof(int64_t) * k_n_state_fields);
    return RS_OK;
}

int rs_state_record_bytes(const rs_model* m, int64_t* bytes) {
    if (!m || !bytes) return fail(RS_E_INVALID, "rs_state_record_bytes: null argument");
    *bytes = record_floats(m) * 4;
    return RS_OK;
}

int rs_model_set_state_layout(rs_model* m, int layout) {
    if (!m || (layout != RS_LAYOUT_PLANES && layout != RS_LAYOUT_ENV_RECORDS))
        return fail(RS_E_INVALID, "rs_model_set_state_layout: bad argument");
    m->layout = layout;
    return RS_OK;
}

int rs_model_get_state_layout(const rs_model* m, int* layout) {
    if (!m || !layout) return fail(RS_E_INVALID, "rs_model_get_state_layout: null argument");
    *layout = m->layout;
    return RS_OK;
}

int rs_state_init(const rs_model* m, void* state, int n_envs, void* stream) {
    if (!m || !state || n_envs < 1) return fail(RS_E_INVALID, "rs_state_init: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int64_t total;
    std::vector<int64_t> offs(k_n_state_fields);
    layout(m, n_envs, offs.data(), &total);
    HIP_OK(hipMemsetAsync(state, 0, total, s));
    DState st = make_state(m, state, n_envs);
    hipLaunchKernelGGL(k_init, env_grid(n_envs), dim3(kBlock), 0, s, m->dm, st);
    HIP_OK(hipGetLastError());
    return RS_OK;
}

int rs_step(const rs_model* m, void* state, int n_envs, int n_substeps, int mode, void* stream) {
    if (!m || !state || n_envs < 1 || n_substeps < 0) return fail(RS_E_INVALID, "rs_step: bad argument");
    if (mode != RS_MODE_PLAIN && mode != RS_MODE_FUSED) return fail(RS_E_INVALID, "rs_step: bad mode");
    hipStream_t s = (hipStream_t)stream;
    DState st = make_state(m, state, n_envs);
    for (int k = 0; k < n_substeps; k++) {
        int r = mode == RS_MODE_PLAIN ? plain_substep(m, st, s) : fused_substep(m, st, s);
        if (r) return r;
    }
    return RS_OK;
}

int rs_stage(const rs_model* m, void* state, int n_envs, int stage, void* stream) {
    if (!m || !state || n_envs < 1 || stage < 0 || stage >= RS_STAGE_COUNT) return fail(RS_E_INVALID, "rs_stage: bad argument");
    DState st = make_state(m, state, n_envs);
    return launch_stage(m, st, stage, (hipStream_t)stream);
}

int rs_forward_kinematics(const rs_model* m, void* state, int n_envs, const int32_t* mask, void* stream) {
    if (!m || !state || n_envs < 1) return fail(RS_E_INVALID, "rs_forward_kinematics: bad argument");
    DState st = make_state(m, state, n_envs);
    hipLaunchKernelGGL(k_fk_masked, env_grid(n_envs), dim3(kBlock), 0, (hipStream_t)stream, m->dm, st, mask);
    HIP_OK(hipGetLastError());
    return RS_OK;
}

int rs_reset_caches(const rs_model* m, void* state, int n_envs, const int32_t* mask, void* stream) {
    if (!m || !state || n_envs < 1) return fail(RS_E_INVALID, "rs_reset_caches: bad argument");
    DState st = make_state(m, state, n_envs);
    hipLaunchKernelGGL(k_reset_caches, env_grid(n_envs), dim3(kBlock), 0, (hipStream_t)stream, m->dm, st, mask);
    HIP_OK(hipGetLastError());
    return RS_OK;
}

int rs_errno_reduce(const rs_model* m, const void* state, int n_envs, int32_t* out, void* stream) {
    if (!m || !state || n_envs < 1 || !out) return fail(RS_E_INVALID, "rs_errno_reduce: bad argument");
    hipStream_t s = (hipStream_t)stream;
    DState st = make_state(m, (void*)state, n_envs);
    int32_t* dv = nullptr;
    HIP_OK(hipMallocAsync((void**)&dv, sizeof(int32_t), s));
    HIP_OK(hipMemsetAsync(dv, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_errno_or, dim3((n_envs + 255) / 256), dim3(256), 0, s, st, dv);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, dv, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipFreeAsync(dv, s));
    HIP_OK(hipStreamSynchronize(s));
    return RS_OK;
}

int64_t rs_pair_contact_work_floats(void) { return kPairWorkWords; }

int rs_pair_contact(const rs_model* m, int n, const int32_t* ga, const int32_t* gb, const float* pose_a,
                    const float* pose_b, int method, int lanes, float* work, float* out, int32_t* hit, void* stream) {
    if (!m || n < 0 || (n > 0 && (!ga || !gb || !pose_a || !pose_b || !out || !hit)))
        return fail(RS_E_INVALID, "rs_pair_contact: bad argument");
    if (method != RS_PAIR_GJK && method != RS_PAIR_MPR) return fail(RS_E_INVALID, "rs_pair_contact: bad method");
    if (lanes != 1 && lanes != kWave) return fail(RS_E_INVALID, "rs_pair_contact: lanes must be 1 or 64");
    if (lanes == kWave && method != RS_PAIR_GJK)
        return fail(RS_E_INVALID, "rs_pair_contact: the wavefront detection is GJK/EPA only");
    if (method == RS_PAIR_GJK && n > 0 && !work) return fail(RS_E_INVALID, "rs_pair_contact: GJK/EPA needs a work buffer");
    if (n == 0) return RS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (lanes == kWave)
        hipLaunchKernelGGL(k_pair_gjk_wave, dim3((unsigned)n), dim3(kWave), 0, s, m->dm, n, ga, gb, pose_a, pose_b,
                           work, out, hit);
    else
        hipLaunchKernelGGL(k_pair_lane, env_grid(n), dim3(kBlock), 0, s, m->dm, n, ga, gb, pose_a, pose_b, method,
                           work, out, hit);
    HIP_OK(hipGetLastError());
    return RS_OK;
}

// Device address of `scratch` on each device, taken once per device with that device current (the array's name in
// host code is only a stand-in for the symbol, not a device address).
static constexpr int kMaxDevices = 64;
static int32_t* g_scratch_dev[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_of_current_device(int32_t** out) {
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(RS_E_DEVICE, "rs_errno_reduce_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(RS_E_DEVICE, "rs_errno_reduce_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = (int32_t*)p;
    }
    *out = g_scratch_dev[dev];
    return RS_OK;
}

int rs_errno_reduce_noalloc(const rs_model* m, const void* state, int n_envs, int32_t* out, void* stream) {
    if (!m || !state || n_envs < 1 || !out) return fail(RS_E_INVALID, "rs_errno_reduce_noalloc: bad argument");
    const int64_t n_blocks = ((int64_t)n_envs + kErrnoBlock - 1) / kErrnoBlock;
    if (n_blocks > kErrnoMaxPartials)
        return fail(RS_E_INVALID, "rs_errno_reduce_noalloc: n_envs exceeds the static scratch (at most " +
                                      std::to_string((int64_t)kErrnoMaxPartials * kErrnoBlock) + " environments)");
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev != m->device) return fail(RS_E_INVALID, "rs_errno_reduce_noalloc: the model lives on another device");
    int32_t* sp = nullptr;
    const int r = scratch_of_current_device(&sp);
    if (r != RS_OK) return r;
    hipStream_t s = (hipStream_t)stream;
    const DState st = make_state(m, (void*)state, n_envs);
    hipLaunchKernelGGL(k_errno_partial, dim3((unsigned)n_blocks), dim3(kErrnoBlock), 0, s, st, sp);
    hipLaunchKernelGGL(k_errno_final, dim3(1), dim3(kErrnoBlock), 0, s, (const int32_t*)sp, (int)n_blocks,
                       sp + kErrnoMaxPartials);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, sp + kErrnoMaxPartials, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return RS_OK;
}

int rs_graph_create(const rs_model* m, void* state, int n_envs, int n_substeps, int mode, void* stream, rs_graph** out) {
    if (!m || !state || !out || n_envs < 1 || n_substeps < 1) return fail(RS_E_INVALID, "rs_graph_create: bad argument");
    hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(RS_E_INVALID, "rs_graph_create: capture needs a non-default stream");
    rs_graph* g = new rs_graph();
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = rs_step(m, state, n_envs, n_substeps, mode, stream);
    hipError_t e = hipStreamEndCapture(s, &g->graph);
    if (r != RS_OK || e != hipSuccess) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return r != RS_OK ? r : fail(RS_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g->graph);
        delete g;
        return fail(RS_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    *out = g;
    return RS_OK;
}

int rs_graph_launch(rs_graph* g, void* stream) {
    if (!g) return fail(RS_E_INVALID, "rs_graph_launch: null graph");
    HIP_OK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return RS_OK;
}

int rs_batch_create(int max_segments, rs_batch** out) {
    if (max_segments < 1 || !out) return fail(RS_E_INVALID, "rs_batch_create: bad argument");
    rs_batch* bt = new rs_batch();
    bt->max_segments = max_segments;
    for (int i = 0; i < rs_batch::kRing; i++) {
        if (hipHostMalloc((void**)&bt->host[i], sizeof(FusedSegment) * max_segments, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void**)&bt->dev[i], sizeof(FusedSegment) * max_segments) != hipSuccess ||
            hipEventCreateWithFlags(&bt->done[i], hipEventDisableTiming) != hipSuccess) {
            rs_batch_destroy(bt);
            return fail(RS_E_DEVICE, "rs_batch_create: allocation failed");
        }
    }
    *out = bt;
    return RS_OK;
}

void rs_batch_destroy(rs_batch* bt) {
    if (!bt) return;
    for (int i = 0; i < rs_batch::kRing; i++) {
        if (bt->done[i]) {
            (void)hipEventSynchronize(bt->done[i]);
            (void)hipEventDestroy(bt->done[i]);
        }
        if (bt->dev[i]) (void)hipFree(bt->dev[i]);
        if (bt->host[i]) (void)hipHostFree(bt->host[i]);
    }
    delete bt;
}

int rs_batch_step(rs_batch* bt, int n, const rs_model* const* models, void* const* states, const int* n_envs,
                  void* stream) {
    if (!bt || n < 1 || n > bt->max_segments || !models || !states || !n_envs)
        return fail(RS_E_INVALID, "rs_batch_step: bad argument");
    hipStream_t s = (hipStream_t)stream;
    const int epb = 2;
    int slice_lds = 0, slice_hbm = 0;
    for (int i = 0; i < n; i++) {
        if (!models[i] || !states[i] || n_envs[i] < 1) return fail(RS_E_INVALID, "rs_batch_step: bad segment");
        slice_lds = std::max(slice_lds, lds_slice_floats(models[i]->h, models[i]->fa, true));
        slice_hbm = std::max(slice_hbm, lds_slice_floats(models[i]->h, models[i]->fa, false));
    }
    int max_lds = models[0]->max_lds;
    for (int i = 1; i < n; i++) max_lds = max_lds < models[i]->max_lds ? max_lds : models[i]->max_lds;
    const bool use_lds = (int64_t)epb * slice_lds * 4 <= max_lds;
    const int slice = use_lds ? slice_lds : slice_hbm;
    const int k = bt->slot;
    bt->slot = (bt->slot + 1) % rs_batch::kRing;
    if (bt->pending[k]) HIP_OK(hipEventSynchronize(bt->done[k]));
    FusedSegment* tab = bt->host[k];
    int blocks = 0;
    for (int i = 0; i < n; i++) {
        tab[i].m = models[i]->dm;
        tab[i].st = make_state(models[i], states[i], n_envs[i]);
        tab[i].fa = models[i]->fa;
        set_lds_slice(models[i]->h, &tab[i].fa, use_lds, slice);
        tab[i].fa.epb = epb;
        tab[i].fa.sched = 0;
        tab[i].first_block = blocks;
        tab[i].n_blocks = (n_envs[i] + epb - 1) / epb;
        blocks += tab[i].n_blocks;
    }
    HIP_OK(hipMemcpyAsync(bt->dev[k], tab, sizeof(FusedSegment) * n, hipMemcpyHostToDevice, s));
    HIP_OK(hipEventRecord(bt->done[k], s));
    bt->pending[k] = true;
    const dim3 grid((unsigned)blocks), block((unsigned)(kWave * epb));
    launch_fused_multi(use_lds, grid, block, (size_t)epb * slice * 4, s, bt->dev[k], n);
    HIP_OK(hipGetLastError());
    return RS_OK;
}

void rs_graph_destroy(rs_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

}  // extern "C"
