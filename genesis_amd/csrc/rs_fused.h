// Fused substep: the whole of RigidSolver.substep (rigid_solver.py:912-980, 2719-2891) for one environment per
// 64-lane wavefront, in a single launch. The environment's small fields are staged into a private region (LDS when
// the model fits, else the env-major work field) and written back to the [row][env] planes at the end; see
// rs_wave.h for the execution model. Results equal the per-stage kernels: the same arithmetic in the same order.
#pragma once
#include "rs_collision.h"
#include "rs_constraint.h"
#include "rs_dyn.h"
#include "rs_wave.h"
#include "rs_wave_dyn.h"
#include "rs_wave_solver.h"

namespace rs {

// broadphase.py:141-396 on one wavefront, producing the same pair list in the same order as broadphase_sap.
// The endpoints are sorted by rank (rank = number of smaller values + number of equal values earlier in the previous
// order, i.e. the stable sort that the insertion sort computes); NaN values fall back to the insertion sort itself.
// In the sequential sweep, the active list when the min endpoint at sorted position p is reached holds exactly the
// geoms whose min position is below p and whose max position is above p, in min-position order. So the pair list is
// the list of (a, b), a < b indices into the min endpoints in sorted order, b major and a minor, whose geom a is
// still active at min endpoint b and which pass the pair filter and the AABB test: for up to 64 geoms each lane
// takes one (a, b) candidate (index t = b (b - 1) / 2 + a) and a ballot places the survivors in that order; larger
// models let lane p walk its own candidates. The active_buffer field holds each geom's max position here (it is
// scratch in both paths).
__device__ inline void wave_broadphase(const DModel& m, const LState& loc, const DState& st, int b, int lane) {
    const int G = m.G, N = 2 * G;
    if (FP(loc, first_time)[0]) {
        sap_init_endpoints(m, loc, 0);
    } else {
        for (int i = lane; i < N; i += kWave) {
            const int g = FP(loc, sort_i_g)[i];
            FP(loc, sort_value)[i] = FP(loc, sort_is_max)[i] ? FP(loc, aabb_max)[3 * g] : FP(loc, aabb_min)[3 * g];
        }
    }
    wave_sync();
    bool nan = false;
    for (int i = lane; i < N; i += kWave) nan |= isnan(FP(loc, sort_value)[i]);
    if (wave_ballot(nan) != 0 || N > 2 * kWave) {
        sap_insertion_sort(loc, 0, N);
    } else {
        float v[2];
        int ig[2], mx[2], rank[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = lane + h * kWave;
            rank[h] = -1;
            if (i < N) {
                v[h] = FP(loc, sort_value)[i];
                ig[h] = FP(loc, sort_i_g)[i];
                mx[h] = FP(loc, sort_is_max)[i];
                int r = 0;
                for (int j = 0; j < N; j++) {
                    const float vj = FP(loc, sort_value)[j];
                    r += (vj < v[h]) || (vj == v[h] && j < i);
                }
                rank[h] = r;
            }
        }
        wave_sync();
#pragma unroll
        for (int h = 0; h < 2; h++)
            if (rank[h] >= 0) {
                FP(loc, sort_value)[rank[h]] = v[h];
                FP(loc, sort_i_g)[rank[h]] = ig[h];
                FP(loc, sort_is_max)[rank[h]] = mx[h];
            }
    }
    wave_sync();
    for (int p = lane; p < N; p += kWave)
        if (FP(loc, sort_is_max)[p]) FP(loc, active_buffer)[FP(loc, sort_i_g)[p]] = p;
    wave_sync();
    int placed = 0;
    if (G <= kWave) {
        // lane k < G: sorted position and geom of the k-th min endpoint (the k-th set bit of the min-endpoint mask)
        const bool lo_min = lane < N && !FP(loc, sort_is_max)[lane];
        const bool hi_min = lane + kWave < N && !FP(loc, sort_is_max)[lane + kWave];
        const uint64_t m_lo = wave_ballot(lo_min), m_hi = wave_ballot(hi_min);
        const int n_lo = __popcll(m_lo);
        int pos_k = 0, g_k = 0;
        if (lane < G) {
            const uint64_t mask = lane < n_lo ? m_lo : m_hi;
            const int k = lane < n_lo ? lane : lane - n_lo;
            int lo = 0, hi = kWave - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (__popcll(mask & ((2ull << mid) - 1)) > k) hi = mid;
                else lo = mid + 1;
            }
            pos_k = (lane < n_lo ? 0 : kWave) + lo;
            g_k = FP(loc, sort_i_g)[pos_k];
        }
        const int T = G * (G - 1) / 2;
        for (int t0 = 0; t0 < T; t0 += kWave) {
            const int t = t0 + lane;
            int bb = (int)((1.f + sqrtf(1.f + 8.f * (float)t)) * 0.5f);
            while (bb > 1 && bb * (bb - 1) / 2 > t) bb--;
            while ((bb + 1) * bb / 2 <= t) bb++;
            const int aa = t - bb * (bb - 1) / 2;
            const int src_a = t < T ? aa : 0, src_b = t < T ? bb : 0;
            const int gA = __shfl(g_k, src_a, kWave), gB = __shfl(g_k, src_b, kWave);
            const int pB = __shfl(pos_k, src_b, kWave);
            bool emit = false;
            int ga = 0, gb = 0;
            if (t < T && FP(loc, active_buffer)[gA] > pB) {
                ga = min(gA, gB);
                gb = max(gA, gB);
                const int pidx = m.collision_pair_idx[ga * G + gb];
                if (pidx != -1) {
                    if (aabb_overlap(loc, 0, ga, gb)) emit = true;
                    else st3(st.cache_normal, pidx, st.B, b, zero3());
                }
            }
            const uint64_t em = wave_ballot(emit);
            if (emit) {
                const int idx = placed + __popcll(em & ((1ull << lane) - 1));
                if (idx < m.BROAD) {
                    st.broad_pairs[(size_t)(2 * idx) * st.B + b] = ga;
                    st.broad_pairs[(size_t)(2 * idx + 1) * st.B + b] = gb;
                }
            }
            placed += __popcll(em);
        }
    }
    for (int p0 = 0; G > kWave && p0 < N; p0 += kWave) {
        const int p = p0 + lane;
        const int gi = (p < N && !FP(loc, sort_is_max)[p]) ? FP(loc, sort_i_g)[p] : -1;
        int cnt = 0;
        for (int pass = 0; pass < 2; pass++) {
            int first = 0;
            if (pass == 1) {
                int total;
                first = placed + wave_excl_scan(cnt, lane, total);
                placed += total;
            }
            int k = 0;
            for (int q = 0; gi >= 0 && q < p; q++) {
                if (FP(loc, sort_is_max)[q]) continue;
                const int g = FP(loc, sort_i_g)[q];
                if (FP(loc, active_buffer)[g] < p) continue;  // its max endpoint came first: no longer active
                const int ga = min(g, gi), gb = max(g, gi);
                const int pidx = m.collision_pair_idx[ga * G + gb];
                if (pidx == -1) continue;
                if (!aabb_overlap(loc, 0, ga, gb)) {
                    if (pass == 0) st3(st.cache_normal, pidx, st.B, b, zero3());
                    continue;
                }
                if (pass == 0) {
                    cnt++;
                } else {
                    const int idx = first + k++;
                    if (idx < m.BROAD) {
                        st.broad_pairs[(size_t)(2 * idx) * st.B + b] = ga;
                        st.broad_pairs[(size_t)(2 * idx + 1) * st.B + b] = gb;
                    }
                }
            }
        }
    }
    if (placed > m.BROAD) FP(loc, errbits)[0] = FP(loc, errbits)[0] | RS_ERR_OVERFLOW_CANDIDATE_CONTACTS;
    FP(loc, n_broad)[0] = min(placed, m.BROAD);
    wave_sync();
}

// narrowphase.py:1988-2189 on one wavefront. Lanes take broadphase pairs; a pair's contacts go to its lane's slot
// array (capacity n_contacts_per_pair) and are then compacted into the contact rows with a prefix sum, in pair order,
// convex pairs first and plane-box pairs second - the order in which the sequential narrowphase appends them. The
// four perturbed detections of each multi-contact pair that collided run one per lane (each lane takes a (pair,
// detection) item); the pair's lane then merges their results in detection order, as the sequential loop does.
// mc: per-pair records (kMcRec floats per lane), the item map and the item results, in the work field. Detections
// handed to GJK/EPA run on the whole wavefront when there is one (polytope in gjk_lds, LDS or nullptr; in gjk_mem when
// it outgrows that) and one per lane when there are several (polytopes in the lane slots at gjk_mem). Returns the
// GJK/EPA work it ran (uniform): lane passes << 16 | single wavefront detections, the scheduler's cost hint.
__device__ __forceinline__ int wave_narrowphase(const DModel& m, const LState& loc, const DState& st, int b, float* pc,
                                       float* mc, float* gjk_lds, float* gjk_mem, int lane) {
    int gjk_passes = 0, gjk_singles = 0;
    const int nb = FP(loc, n_broad)[0];
    int placed = 0;  // uniform: contacts produced so far
    SlotSink sink{pc + (size_t)lane * m.NCPP * 8, m.NCPP, 0};
    float* rec = mc;
    int* imap = (int*)(mc + kWave * kMcRec);
    float* res = mc + kWave * kMcRec + 4 * kWave;
    for (int pass = 0; pass < 2; pass++) {
        for (int p0 = 0; p0 < nb; p0 += kWave) {
            const int p = p0 + lane;
            sink.n = 0;
            int ga = 0, gb = 0, pair = 0;
            ConvexFirst f;
            f.perturb = false;
            bool conv = false;
            if (p < nb) {
                ga = st.broad_pairs[(size_t)(2 * p) * st.B + b];
                gb = st.broad_pairs[(size_t)(2 * p + 1) * st.B + b];
                if (m.geom_type[ga] > m.geom_type[gb]) {
                    int t = ga;
                    ga = gb;
                    gb = t;
                }
                const int ta = m.geom_type[ga], tb = m.geom_type[gb];
                pair = m.collision_pair_idx[min(ga, gb) * m.G + max(ga, gb)];
                if (pass == 0) conv = convex_pair(m, ta, tb, ga, gb) && convex_setup(m, ga, gb, f);
                else if (ta == GT_PLANE && tb == GT_BOX) plane_box(m, loc, 0, ga, gb, sink);
                else if (m.box_box_detection && ta == GT_BOX && tb == GT_BOX) box_box(m, loc, 0, ga, gb, sink);
            }
            if (pass == 0) {
                // round 0: detection 0 of this lane's pair; round 1: the perturbed detections, one (pair, det) item
                // per lane. Both rounds go through the one detection call below.
                int n_items = 0, first_item = 0;
#pragma clang loop unroll(disable)
                for (int round = 0; round < 2; round++) {
                    const int n_it = round == 0 ? 1 : (n_items + kWave - 1) / kWave;
#pragma clang loop unroll(disable)
                    for (int k = 0; k < n_it; k++) {
                        const int it = k * kWave + lane;
                        bool act = round == 0 && conv;
                        int dga = ga, dgb = gb, det = 0;
                        ConvexFirst g = f;
                        if (round == 1 && it < n_items) {
                            act = true;
                            const int o = imap[it] >> 2;
                            det = (imap[it] & 3) + 1;
                            const float* r = rec + o * kMcRec;
                            g.n0 = mk3(r[0], r[1], r[2]);
                            g.p0 = mk3(r[3], r[4], r[5]);
                            g.ax0 = mk3(r[6], r[7], r[8]);
                            g.ax1 = mk3(r[9], r[10], r[11]);
                            g.pair = ((const int*)r)[12];
                            dga = ((const int*)r)[13];
                            dgb = ((const int*)r)[14];
                            g.tolerance = r[15];
                        }
                        f3 n = zero3(), cp = zero3();
                        float pen = 0.f;
                        f4q qr{0.f, 0.f, 0.f, 0.f};
                        bool col = false, want_gjk = false;
                        f3 pa = zero3(), pb = zero3();
                        f4q qa = qident(), qb = qident();
                        if (act) {
                            convex_pose(m, loc, 0, dga, dgb, g, det, pa, qa, pb, qb, qr);
                            col = convex_detect(m, st, b, dga, dgb, g.pair, det == 0, g.tolerance, pa, qa, pb, qb, n,
                                                cp, pen, want_gjk);
                        }
                        // the detections handed to GJK/EPA: a single one runs on the whole wavefront (polytope in
                        // LDS); several run one per lane, each with its polytope in the lane's slot of the work
                        // field; the lane that owns a detection takes its result
                        uint64_t gm = wave_ballot(act && want_gjk);
                        if (gm && !(gm & (gm - 1))) {
                            gjk_singles++;
                            const int l = __builtin_ctzll(gm);
                            const GjkPair gp = make_gjk_pair(m, rl_i(dga, l), rl_i(dgb, l), rl3(pa, l), rl4(qa, l),
                                                             rl3(pb, l), rl4(qb, l));
                            f3 gn = zero3(), gpos = zero3();
                            float gpen = 0.f;
                            const bool gcol = gjk_contact_wave(m, gp, lane, gjk_lds, gjk_mem, gn, gpos, gpen);
                            if (lane == l) col = apply_gjk(gcol, gn, gpos, gpen, n, cp, pen);
                        } else if (gm) {
                            gjk_passes++;
                            const bool mine = (gm >> lane) & 1;
                            const GjkPair gp = make_gjk_pair(m, dga, dgb, pa, qa, pb, qb);
                            f3 gn = zero3(), gpos = zero3();
                            float gpen = 0.f;
                            bool gcol = false;
                            gjk_contact_lanes(m, gp, mine, lane, gjk_mem, gcol, gn, gpos, gpen);
                            if (mine) col = apply_gjk(gcol, gn, gpos, gpen, n, cp, pen);
                        }
                        if (round == 0) {
                            if (act) convex_first(m, loc, 0, st, b, ga, gb, col, n, cp, pen, sink, f);
                        } else if (act) {
                            if (col) convex_correct(m, g, qr, n, cp, pen);
                            float* q = res + (size_t)it * kMcRes;
                            q[0] = n.x; q[1] = n.y; q[2] = n.z;
                            q[3] = cp.x; q[4] = cp.y; q[5] = cp.z;
                            q[6] = pen;
                            q[7] = col ? 1.f : 0.f;
                        }
                    }
                    if (round == 0) {
                        if (!wave_ballot(f.perturb)) break;
                        first_item = wave_excl_scan(f.perturb ? 4 : 0, lane, n_items);
                        if (f.perturb) {
                            float* r = rec + lane * kMcRec;
                            r[0] = f.n0.x; r[1] = f.n0.y; r[2] = f.n0.z;
                            r[3] = f.p0.x; r[4] = f.p0.y; r[5] = f.p0.z;
                            r[6] = f.ax0.x; r[7] = f.ax0.y; r[8] = f.ax0.z;
                            r[9] = f.ax1.x; r[10] = f.ax1.y; r[11] = f.ax1.z;
                            ((int*)r)[12] = f.pair;
                            ((int*)r)[13] = ga;
                            ((int*)r)[14] = gb;
                            r[15] = f.tolerance;
                            for (int j = 0; j < 4; j++) imap[first_item + j] = lane * 4 + j;
                        }
                    }
                    wave_sync();
                }
                if (f.perturb) {
                    int n_con = 1;
                    for (int j = 0; j < 4; j++) {
                        const float* q = res + (size_t)(first_item + j) * kMcRes;
                        if (q[7] != 0.f)
                            convex_merge(sink, n_con, ga, gb, f, mk3(q[0], q[1], q[2]), mk3(q[3], q[4], q[5]), q[6]);
                    }
                }
                wave_sync();
            }
            int total;
            const int first = placed + wave_excl_scan(sink.n, lane, total);
            for (int k = 0; k < sink.n; k++) {
                const int i = first + k;
                if (i < m.KMAX) {
                    const float* s = sink.slot + 8 * k;
                    write_contact(m, st, b, i, ga, gb, mk3(s[0], s[1], s[2]), mk3(s[3], s[4], s[5]), s[6], pair);
                }
            }
            placed += total;
        }
    }
    if (placed > m.KMAX) FP(loc, errbits)[0] = FP(loc, errbits)[0] | RS_ERR_OVERFLOW_COLLISION_PAIRS;
    FP(loc, n_contacts)[0] = min(placed, m.KMAX);
    wave_sync();
    return min(gjk_passes, 0x7fff) << 16 | min(gjk_singles, 0xffff);
}

// dev_integrate (rigid_solver.py kernel_step_2's semi-implicit Euler) with its dofs, links and coordinates split over
// the lanes: every entry is computed by the same arithmetic as there, so the result is the same to the bit; the NaN
// check is a ballot over the lanes.
__device__ inline void wave_integrate(const DModel& m, const LState& loc, int lane) {
    const float dt = m.dt;
    for (int d = lane; d < m.D; d += kWave) FP(loc, vel_next)[d] = FP(loc, vel)[d] + FP(loc, acc)[d] * dt;
    wave_sync();
    for (int l = lane; l < m.L; l += kWave) {
        const int ds = m.link_dof_start[l];
        if (m.link_dof_end[l] == ds) continue;
        const int qs = m.link_q_start[l], qe = m.link_q_end[l];
        const int jt = m.joint_type[m.link_joint_start[l]];
        const float* q = FP(loc, qpos);
        const float* v = FP(loc, vel_next);
        float* qn = FP(loc, qpos_next);
        if (jt == JT_FREE)
            for (int k = 0; k < 3; k++) qn[qs + k] = q[qs + k] + v[ds + k] * dt;
        if (jt == JT_SPHERICAL || jt == JT_FREE) {
            const int ro = jt == JT_FREE ? 3 : 0;
            f4q r0{q[qs + ro], q[qs + ro + 1], q[qs + ro + 2], q[qs + ro + 3]};
            f3 ang = mk3(v[ds + ro], v[ds + ro + 1], v[ds + ro + 2]) * dt;
            f4q r = qtrans(rotvec_quat(ang, m.eps), r0);
            qn[qs + ro] = r.w;
            qn[qs + ro + 1] = r.x;
            qn[qs + ro + 2] = r.y;
            qn[qs + ro + 3] = r.z;
        } else {
            for (int k = 0; k < qe - qs; k++) qn[qs + k] = q[qs + k] + v[ds + k] * dt;
        }
    }
    wave_sync();
    bool bad = false;
    for (int d = lane; d < m.D; d += kWave) bad = bad || isnan(FP(loc, vel_next)[d]);
    for (int i = lane; i < m.Q; i += kWave) bad = bad || isnan(FP(loc, qpos_next)[i]);
    if (wave_ballot(bad) == 0) {
        for (int d = lane; d < m.D; d += kWave) FP(loc, vel)[d] = FP(loc, vel_next)[d];
        for (int i = lane; i < m.Q; i += kWave) FP(loc, qpos)[i] = FP(loc, qpos_next)[i];
    } else {
        FP(loc, errbits)[0] = FP(loc, errbits)[0] | RS_ERR_INVALID_ACC_NAN;
    }
    wave_sync();
}

// One environment's substep on the calling wavefront (w = wavefront index in the workgroup).
// kPath: 1 = models in the first row-tile class (at most kTileMaxDofs dofs), 4 = the second (up to kTileWideDofs),
// 0 = the others, 2 = any (decided per call: the shared launch of several models). Returns the narrowphase's GJK/EPA
// work (see wave_narrowphase).
template <bool kLds, int kPath = 2>
__device__ __forceinline__ int fused_env(const DModel& m, const DState& st, const FusedArgs& fa, int b, int w,
                                          float* lds_env) {
    const int lane = lane_id();
    float* wk = st.wk + (size_t)b * fa.work_stride;
    float* slice = lds_env + (size_t)w * fa.lds_wave;
    // With one record per environment (st.LS > 0) the record holds the local rows word for word as the private region
    // lays them out, so a launch that keeps the private region in global memory works on the record in place instead
    // of copying it into the work field and back (for 120 dofs that copy is three 120 x 120 matrices each way).
    const bool in_place = !kLds && st.LS > 0;
    float* base = kLds ? slice : in_place ? (float*)((uint32_t*)(void*)st.qpos + (size_t)b * st.LS) : wk + fa.off_loc;
    float* tile = slice + fa.tile_off;
    const LState loc = make_local(base, fa);
    const WorkRows R{wk + fa.off_jac, wk + fa.off_rowf, wk + fa.off_tmp, m.CMAX, m.D};

    if (!in_place) stage_local<true>(st, (uint32_t*)base, fa.n_local, b, lane);
    wave_sync();
    // An environment that had many constraint rows in its last substep is likely the longest of this one: it takes
    // issue priority over the other wavefronts of its SIMD, which shortens the step's tail (row-tile models, whose
    // steps are bound by their few heavy environments; scheduling only, no result depends on it).
    if constexpr (kPath == 1 || kPath == 4)
        if (FP(loc, n_con)[0] >= kSchedHeavyRows) __builtin_amdgcn_s_setprio(3);
    // kernel_step_1: CRBA, L^T D L, actuation + passive + bias forces, acc_smooth
    wave_mass_matrix(m, loc, lane);
    wave_factor_mass(m, loc, lane);
    wave_torque_passive(m, loc, lane);
    wave_update_acc(m, loc, lane, false);
    wave_rne_and_qacc(m, loc, lane);
    // Collider.detection
    int gjk_work = 0;
    if (m.enable_collision) {
        for (int g = lane; g < m.G; g += kWave) dev_aabb(m, loc, 0, g);
        wave_sync();
        if (m.P > 0) {
            const int n_old = FP(loc, n_contacts)[0];
            for (int i = lane; i < n_old; i += kWave) clear_contacts(st, b, i, i + 1);
            wave_broadphase(m, loc, st, b, lane);
            gjk_work = wave_narrowphase(m, loc, st, b, wk + fa.off_pc, wk + fa.off_mc,
                                        fa.gjk_off >= 0 ? slice + fa.gjk_off : nullptr, wk + fa.off_gjk, lane);
        }
    }
    // ConstraintSolver.resolve
    if (!m.disable_constraint) {
        wave_constraints(m, loc, st, b, R, lane, fa.tiled ? tile : nullptr);
        if constexpr (kPath == 1) wave_solve<kTileMaxDofs>(m, loc, st, b, R, lane, tile);
        else if constexpr (kPath == 4) wave_solve<kTileWideDofs>(m, loc, st, b, R, lane, tile);
        else if constexpr (kPath == 0) wave_solve<0>(m, loc, st, b, R, lane, tile);
        else if (fa.tiled == kTileMaxDofs) wave_solve<kTileMaxDofs>(m, loc, st, b, R, lane, tile);
        else if (fa.tiled == kTileWideDofs) wave_solve<kTileWideDofs>(m, loc, st, b, R, lane, tile);
        else wave_solve<0>(m, loc, st, b, R, lane, tile);
    }
    // kernel_step_2: cacc, semi-implicit Euler, forward kinematics and velocities
    wave_update_acc(m, loc, lane, true);
    if (m.integrator != RS_INTEGRATOR_APPROXIMATE_IMPLICITFAST) dev_implicit_damping(m, loc, 0);
    wave_integrate(m, loc, lane);
    wave_kinematics(m, loc, lane);
    if (!in_place) stage_local<false>(st, (uint32_t*)base, fa.n_local, b, lane);
    return gjk_work;
}

// Consecutive environments go to one XCD (workgroups are dealt round-robin over the 8 XCDs), so that the planes'
// cache lines, shared by neighbouring environments, stay in one L2.
__device__ __forceinline__ int xcd_block(int blk, int nblk) {
    return (nblk & 7) == 0 ? (blk & 7) * (nblk >> 3) + (blk >> 3) : blk;
}

// flips the schedule parity and empties the queues the coming substep appends to (one thread, before each substep)
static __global__ void k_sched_flip(DState st) {
    int* ctl = st.sched + (size_t)2 * kSchedClasses * st.B;
    const int par = (ctl[0] & 1) ^ 1;
    ctl[0] = par;
    for (int k = 0; k < kSchedClasses; k++) ctl[1 + (par ^ 1) * kSchedClasses + k] = 0;
}

// kOcc > 0 asks the compiler for at least kOcc wavefronts per SIMD (fewer registers, more spills); 0 leaves the
// register budget to the compiler. The autotuner picks between them. Models above 24 dofs (kPath 0) run at most
// kGenericMaxThreads lanes (4 environments) per workgroup, so that a wavefront may hold all 512 registers of its SIMD
// (the matrix-core Hessian keeps its 36 tile accumulators resident).
template <bool kLds, int kOcc, int kPath>
__global__ void __launch_bounds__(kPath == 0 ? kGenericMaxThreads : 512)
__attribute__((amdgpu_waves_per_eu(kOcc > 0 ? kOcc : 1, 8)))
k_fused_wave(DModel m, DState st, FusedArgs fa) {
    extern __shared__ float lds_env[];
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int b = xcd_block((int)blockIdx.x, (int)gridDim.x) * fa.epb + w;
    int* ctl = nullptr;
    int par = 0;
    if (fa.sched) {
        ctl = st.sched + (size_t)2 * kSchedClasses * st.B;
        par = ctl[0] & 1;
        int cnt[kSchedClasses], total = 0;
        bool valid = true;
#pragma unroll
        for (int k = 0; k < kSchedClasses; k++) {
            cnt[k] = ctl[1 + par * kSchedClasses + k];
            valid &= cnt[k] >= 0 && cnt[k] <= st.B;
            total += cnt[k];
        }
        int pos = (int)blockIdx.x * fa.epb + w;  // dispatch order
        if (valid && total == st.B && pos < st.B) {
            int k = 0;
            while (k < kSchedClasses - 1 && pos >= cnt[k]) pos -= cnt[k++];
            const int e = __builtin_amdgcn_readfirstlane(st.sched[(size_t)(par * kSchedClasses + k) * st.B + pos]);
            if (e >= 0 && e < st.B) b = e;
        }
    }
    if (b >= st.B) return;  // whole wavefronts only
    const int gjk_work = fused_env<kLds, kPath>(m, st, fa, b, w, lds_env);
    if (fa.sched && lane_id() == 0) {
        const int cls = sched_class(FV(st, n_con, b)[0], gjk_work);
        const int q = par ^ 1;
        const int slot = atomicAdd(&ctl[1 + q * kSchedClasses + cls], 1);
        if (slot < st.B) st.sched[(size_t)(q * kSchedClasses + cls) * st.B + slot] = b;
    }
}

// One batch of a shared launch: a model, its state slab, and the first workgroup that serves it.
struct FusedSegment {
    DModel m;
    DState st;
    FusedArgs fa;
    int first_block, n_blocks;
};

// Several models and slabs in one launch (the serving loop's co-batched step). Every segment uses the launch's
// environments-per-workgroup; a workgroup finds its segment by scanning the table (a few dozen entries at most).
template <bool kLds>
__global__ void __launch_bounds__(512) k_fused_multi(const FusedSegment* __restrict__ segs, int n_segs) {
    extern __shared__ float lds_env[];
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int blk = xcd_block((int)blockIdx.x, (int)gridDim.x);
    int s = 0;
    while (s + 1 < n_segs && segs[s + 1].first_block <= blk) s++;
    const FusedSegment& g = segs[s];
    const int b = (blk - g.first_block) * g.fa.epb + w;
    if (b >= g.st.B) return;
    fused_env<kLds>(g.m, g.st, g.fa, b, w, lds_env);
}

}  // namespace rs
