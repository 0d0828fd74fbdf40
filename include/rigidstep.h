/*
 * rigidstep.h - C ABI of the MI355X batched rigid-body step (librigidstep.so).
 *
 * The reference exposes this path only as a Python object (RigidSolver, genesis/engine/solvers/rigid/
 * rigid_solver.py:221) driven by Simulator.step -> RigidSolver.substep (simulator.py:283-286,
 * rigid_solver.py:912-980). These entry points are what a binding of that object needs: build the static model once
 * (RigidSolver.build, rigid_solver.py:317-355), describe/initialise the per-environment state slab
 * (array_class.DataManager, utils/array_class.py:2058), run substeps (substep), run single stages for parity
 * (kernel_step_1/kernel_step_2 pieces, rigid_solver.py:2719-2891; Collider.detection, collider/collider.py:684;
 * ConstraintSolver.resolve, constraint/solver.py:210), and reduce the per-env error bits (check_errno,
 * rigid_solver.py:985-1009).
 *
 * Memory: the caller owns all device memory. The model blob is copied to the device by rs_model_create and owned by
 * the returned handle. The state slab is a caller-allocated device buffer of rs_state_layout().total_bytes bytes whose
 * fields are [rows][n_envs] planes (environment index fastest), float32 or int32, except the per-environment fields,
 * which are per-environment records by default (RS_LAYOUT_ENV_RECORDS below). All work is enqueued on the stream
 * passed by the caller; no entry point synchronises except rs_errno_reduce and the *_sync helpers.
 *
 * Return codes: 0 on success, negative on error (RS_E_*). Error text: rs_last_error().
 */
#ifndef RIGIDSTEP_H
#define RIGIDSTEP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RS_ABI_VERSION 3

enum {
    RS_OK = 0,
    RS_E_INVALID = -1,   /* bad argument */
    RS_E_DEVICE = -2,    /* HIP runtime error */
    RS_E_VERSION = -3,   /* blob / ABI version mismatch */
    RS_E_UNSUPPORTED = -4,
};

/* Per-environment error bits; values equal the reference ErrorCode (utils/array_class.py:91-97). Hibernation is not
 * built, so OVERFLOW_HIBERNATION_ISLANDS is never set; its value is reserved so that the other bits keep theirs. */
enum {
    RS_ERR_OVERFLOW_CANDIDATE_CONTACTS = 1 << 0,
    RS_ERR_OVERFLOW_COLLISION_PAIRS = 1 << 1,
    RS_ERR_OVERFLOW_HIBERNATION_ISLANDS = 1 << 2,
    RS_ERR_INVALID_FORCE_NAN = 1 << 3,
    RS_ERR_INVALID_ACC_NAN = 1 << 4,
};

/* Step implementations. */
enum { RS_MODE_PLAIN = 0, RS_MODE_FUSED = 1 };

/* Stages of one substep, in execution order (see rs_stage). */
enum {
    RS_STAGE_FK = 0,          /* forward kinematics, COM, geom poses, velocities (func_update_cartesian_space +
                                 func_forward_velocity, forward_kinematics.py:1622-1698, :1172-1214) */
    RS_STAGE_CRBA = 1,        /* mass matrix (forward_dynamics.py:292-541) */
    RS_STAGE_FACTOR = 2,      /* L^T D L factorisation (forward_dynamics.py:545-684) */
    RS_STAGE_RNE = 3,         /* actuation, passive, bias forces, acc_smooth (forward_dynamics.py:956-1563) */
    RS_STAGE_AABB = 4,        /* world AABBs (forward_kinematics.py:1297-1320) */
    RS_STAGE_BROADPHASE = 5,  /* sweep-and-prune (collider/broadphase.py:141-396) */
    RS_STAGE_NARROWPHASE = 6, /* contacts (collider/narrowphase.py:1988-2189) */
    RS_STAGE_CONSTRAINTS = 7, /* equality, friction-loss, contact and limit rows (constraint/solver.py:580-1286) */
    RS_STAGE_SOLVE = 8,       /* Newton or CG solve + qacc (constraint/solver.py:2936-3261) */
    RS_STAGE_INTEGRATE = 9,   /* cacc, semi-implicit Euler, NaN guard (rigid_solver.py:2783-2891) */
    RS_STAGE_COUNT = 10,
};

/* ------------------------------------------------------------------------------------------------ model blob */

/* Element count kinds used by the model-array and state-field directories. */
enum {
    RS_N_ONE = 0, RS_N_LINKS, RS_N_JOINTS, RS_N_DOFS, RS_N_QS, RS_N_GEOMS, RS_N_ENTITIES, RS_N_EQUALITIES,
    RS_N_DOFS2, RS_N_GEOMS2, RS_N_PAIRS, RS_N_SUPPORT, RS_N_CONTACTS, RS_N_CONSTRAINTS, RS_N_CONSTRAINTS_DOFS,
    RS_N_BROAD, RS_N_SORT,
    RS_N_WORK,  /* per-environment working set of the fused substep (size from rs_state_field_rows) */
    RS_N_SCHED, /* fused-substep schedule: 2 x 4 cost-class queues of environment ids and one control row */
    RS_N_DOFS_PLANE, /* n_dofs, always a [row][env] plane (state the fused substep keeps outside its private region) */
    RS_N_ONE_PLANE,  /* 1, likewise */
    RS_N_VERTS,      /* collision-geometry vertices (verts_info.init_pos of the reference) */
    RS_N_KINDS
};

#define RS_MAX_MODEL_ARRAYS 96

/* integrators (genesis/constants.py:52-55 integrator) */
#define RS_INTEGRATOR_EULER 0
#define RS_INTEGRATOR_IMPLICITFAST 1
#define RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST 2

/* constraint solvers (genesis/constants.py:59-61 constraint_solver) */
#define RS_SOLVER_CG 0
#define RS_SOLVER_NEWTON 1

typedef struct rs_model_header {
    int32_t abi_version;     /* RS_ABI_VERSION */
    int32_t float_bytes;     /* 4 for the device blob */
    int32_t n_links, n_joints, n_dofs, n_qs, n_geoms, n_entities, n_equalities;
    int32_t n_pairs, n_support, n_verts;
    int32_t max_collision_pairs, max_contact_pairs, max_collision_pairs_broad, n_contacts_per_pair;
    int32_t len_constraints;
    /* options (RigidOptions / SimOptions subset, genesis/options/solvers.py:482-535) */
    float dt;
    float gravity[3];
    int32_t iterations;
    float tolerance;
    int32_t ls_iterations;
    float ls_tolerance;
    int32_t enable_collision, enable_joint_limit, disable_constraint, enable_multi_contact, box_box_detection;
    int32_t integrator, solver_type;                /* RS_INTEGRATOR_*, RS_SOLVER_CG or RS_SOLVER_NEWTON */
    float meaninertia;
    float mc_perturbation, mc_tolerance, mpr_to_gjk_overlap_ratio;
    int32_t n_arrays;                                /* number of directory entries used */
    int64_t array_offset[RS_MAX_MODEL_ARRAYS];       /* byte offset of each array in the blob */
} rs_model_header;

typedef struct rs_array_info {
    const char* name;
    int32_t is_int;   /* 1 = int32, 0 = float */
    int32_t count;    /* RS_N_* kind */
    int32_t width;    /* scalars per element */
} rs_array_info;

/* Directory of model arrays (name, type, element count kind, width) in blob order. */
int rs_model_array_count(void);
int rs_model_array_info(int index, rs_array_info* out);

typedef struct rs_model rs_model;

/* Upload a model blob (layout per the directory above) to the current device. */
int rs_model_create(const rs_model_header* header, const void* blob, size_t blob_bytes, rs_model** out);
void rs_model_destroy(rs_model* model);

/* Overwrite the model's array values with a blob of the same header (counts, options and offsets unchanged), in
 * stream order, then synchronise the stream; the device address, the launch configuration and captured graphs stay
 * valid and see the new values. Used by the dof parameter setters (set_dofs_kp / kv / force_range / damping /
 * stiffness, engine/entities/rigid_entity/rigid_entity.py:3372-3500, which write dofs_info in place). */
int rs_model_update(rs_model* model, const rs_model_header* header, const void* blob, size_t blob_bytes, void* stream);

/* ------------------------------------------------------------------------------------------------ state slab */

/* Directory of per-environment state fields: each is a [rows][n_envs] plane of float32/int32, except fields of kind
 * RS_N_WORK, which are env-major [n_envs][rows] scratch of the fused substep (contents unspecified between calls). */
int rs_state_field_count(void);
int rs_state_field_info(int index, rs_array_info* out);

/* Rows per environment of state field `index` for this model (count of its kind x width, at least 1). */
int rs_state_field_rows(const rs_model* model, int index, int64_t* rows);

/* State layouts of the per-environment (local-kind: one, links, joints, dofs, qs, geoms, entities, dofs^2, SAP
 * endpoint) fields. RS_LAYOUT_PLANES: [rows][n_envs] planes like every other field (Genesis's [n, B] field layout).
 * RS_LAYOUT_ENV_RECORDS (the default): one contiguous record per environment holding all its local rows, so that the
 * fused kernel stages an environment with contiguous copies; element (row r, env b) of local field i is at byte
 * offsets[i] + b * record_bytes + 4 r. Set it before creating the model's slabs: a slab is laid out, and read by
 * every entry point, according to the model's current setting. */
enum { RS_LAYOUT_PLANES = 0, RS_LAYOUT_ENV_RECORDS = 1 };
int rs_model_set_state_layout(rs_model* model, int layout);
int rs_model_get_state_layout(const rs_model* model, int* layout);

/* Byte offset of every state field for n_envs environments, and the slab size. offsets has rs_state_field_count()
 * entries; for local fields in RS_LAYOUT_ENV_RECORDS the offset is inside the record of environment 0. */
int rs_state_layout(const rs_model* model, int n_envs, int64_t* offsets, int64_t* total_bytes);

/* Bytes per environment record of the local fields (0 in RS_LAYOUT_PLANES). */
int rs_state_record_bytes(const rs_model* model, int64_t* record_bytes);

/* Initialise a slab: qpos = qpos0, zero velocities, FORCE control mode, cleared caches, then forward kinematics
 * (Scene.build + reset, scene.py:841-905). */
int rs_state_init(const rs_model* model, void* state, int n_envs, void* stream);

/* Run n_substeps substeps of RigidSolver.substep on every environment. mode: RS_MODE_PLAIN (one kernel per stage,
 * state round-trips through the slab) or RS_MODE_FUSED (one persistent kernel per substep). */
int rs_step(const rs_model* model, void* state, int n_envs, int n_substeps, int mode, void* stream);

/* Run one stage of the plain path (RS_STAGE_*), for per-stage parity checks. */
int rs_stage(const rs_model* model, void* state, int n_envs, int stage, void* stream);

/* Forward kinematics + velocities only on the environments whose mask entry is non-zero (device int32[n_envs]);
 * used after set_qpos / set_dofs_position (rigid_solver.py:2155-2199). mask may be NULL for all environments. */
int rs_forward_kinematics(const rs_model* model, void* state, int n_envs, const int32_t* env_mask, void* stream);

/* Reset warm-start caches (constraint qacc_ws/is_warmstart, collider normal cache, SAP buffer) of masked envs. */
int rs_reset_caches(const rs_model* model, void* state, int n_envs, const int32_t* env_mask, void* stream);

/* Bitwise OR of the per-environment error words into *out (synchronises the stream). */
int rs_errno_reduce(const rs_model* model, const void* state, int n_envs, int32_t* out, void* stream);

/* One contact detection per item, for tests and diagnostics (the detectors of the narrowphase on their own; the
 * restatement's counterpart is the oracle's pair detection): geoms ga[i], gb[i] of the model at world poses
 * pose_a[7 i .. 7 i + 6], pose_b[...] (x, y, z, qw, qx, qy, qz), with GJK/EPA (RS_PAIR_GJK, collider/gjk.py:170-446)
 * or MPR without a warm-start direction (RS_PAIR_MPR, collider/mpr.py:769-818). lanes = 64 runs each item on a whole
 * wavefront (the fused kernel's GJK/EPA for a single detection: polytope in LDS, in work when it outgrows that);
 * lanes = 1 runs one item per lane (GJK/EPA as the fused kernel runs several detections of an environment: a compact
 * polytope per lane in work; MPR as the per-stage kernels run it). GJK/EPA needs work: n *
 * rs_pair_contact_work_floats() floats. Results:
 * out[7 i ..] = normal (from a to b), position, penetration; hit[i] = 1 contact, 0 none, -1 invalid item (geom
 * index out of range, or a plane or terrain geom). Every pointer is a device pointer. */
enum { RS_PAIR_GJK = 0, RS_PAIR_MPR = 1 };
int64_t rs_pair_contact_work_floats(void);
int rs_pair_contact(const rs_model* model, int n, const int32_t* ga, const int32_t* gb, const float* pose_a,
                    const float* pose_b, int method, int lanes, float* work, float* out, int32_t* hit, void* stream);

/* rs_errno_reduce without any allocation. No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture of the surrounding loop, tight serving loops). Its partial results live in a
 * static device array of the library (one per GPU, held until the process ends), which limits n_envs to 1,048,576
 * per call (larger counts are rejected with RS_E_INVALID). Calls that overlap on one GPU share that array and race.
 * The model must live on the current device. Synchronises the stream, like rs_errno_reduce. */
int rs_errno_reduce_noalloc(const rs_model* model, const void* state, int n_envs, int32_t* out, void* stream);

/* Launch configuration of the fused substep for one model: one 64-lane wavefront per environment,
 * envs_per_block wavefronts per workgroup (1..8), whether the environment's small fields are staged in LDS
 * (use_lds = 1, only when they fit) or in the env-major work field, and whether environments are started in order of
 * their previous substep's cost (schedule = 1; every environment's result is the same in any order), and the kernel's
 * register budget (occupancy 0, 3 or 4; models with at most 16 dofs have all three, models with 17 to 24 dofs 0 and 3,
 * larger models 0 only). rs_fused_config_set rejects configurations the device cannot run. The model keeps the
 * configuration for every later RS_MODE_FUSED step. */
typedef struct rs_fused_config {
    int32_t envs_per_block;
    int32_t use_lds;
    int32_t lds_bytes_per_env;  /* out: private-region size of one environment */
    int32_t schedule;           /* 1: start the environments that were costliest in the previous substep first */
    int32_t occupancy;          /* 0: compiler's register budget; 3 or 4: kernel built for that many wavefronts per SIMD */
    int32_t gjk_lds;            /* 1: GJK/EPA polytope of single detections in LDS (models with convex mesh geoms;
                                   default 0: in the work field, leaving LDS for more wavefronts per compute unit) */
} rs_fused_config;
int rs_fused_config_get(const rs_model* model, rs_fused_config* out);
int rs_fused_config_set(rs_model* model, const rs_fused_config* config);

/* Shared launches for serving: one fused substep of n independent batches - each a model, its state slab and its
 * environment count, models may differ - in a single kernel launch. Each environment's result is the same as when
 * its slab is stepped alone with rs_step(RS_MODE_FUSED). The batch object owns a ring of four small pinned/device
 * segment tables (up to max_segments entries each); rs_batch_step waits only for the table copy it issued four calls
 * earlier before reusing that table. No counterpart in the reference (its scenes step one at a time). */
typedef struct rs_batch rs_batch;
int rs_batch_create(int max_segments, rs_batch** out);
void rs_batch_destroy(rs_batch* batch);
int rs_batch_step(rs_batch* batch, int n, const rs_model* const* models, void* const* states, const int* n_envs,
                  void* stream);

/* hipGraph capture of n_substeps fused substeps on a stream; replay with rs_graph_launch. */
typedef struct rs_graph rs_graph;
int rs_graph_create(const rs_model* model, void* state, int n_envs, int n_substeps, int mode, void* stream,
                    rs_graph** out);
int rs_graph_launch(rs_graph* graph, void* stream);
void rs_graph_destroy(rs_graph* graph);

const char* rs_last_error(void);
int rs_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RIGIDSTEP_H */
