// Fused substep: the whole of RigidSolver.substep (rigid_solver.py:912-980, 2719-2891) for one environment per
// 64-lane wavefront, in a single launch. The environment's small fields are staged into a private region (LDS when
// the model fits, else the env-major work field) and written back to the [row][env] planes at the end; see
// rs_wave.h for the execution model. Results equal the per-stage kernels: the same arithmetic in the same order.
#pragma once
#include "rs_collision.h"
#include "rs_constraint.h"
#include "rs_dyn.h"
#include "rs_wave.h"
#include "rs_wave_solver.h"

namespace rs {

// narrowphase.py:1988-2189 on one wavefront. Lanes take broadphase pairs; a pair's contacts go to its lane's slot
// array (capacity n_contacts_per_pair) and are then compacted into the contact rows with a prefix sum, in pair order,
// convex pairs first and plane-box pairs second - the order in which the sequential narrowphase appends them.
__device__ inline void wave_narrowphase(const DModel& m, const DState& loc, const DState& st, int b, float* pc,
                                        int lane) {
    const int nb = loc.n_broad[0];
    int placed = 0;  // uniform: contacts produced so far
    SlotSink sink{pc + (size_t)lane * m.NCPP * 8, m.NCPP, 0};
    for (int pass = 0; pass < 2; pass++) {
        for (int p0 = 0; p0 < nb; p0 += kWave) {
            const int p = p0 + lane;
            sink.n = 0;
            int ga = 0, gb = 0, pair = 0;
            if (p < nb) {
                ga = loc.broad_pairs[2 * p];
                gb = loc.broad_pairs[2 * p + 1];
                if (m.geom_type[ga] > m.geom_type[gb]) {
                    int t = ga;
                    ga = gb;
                    gb = t;
                }
                const int ta = m.geom_type[ga], tb = m.geom_type[gb];
                pair = m.collision_pair_idx[min(ga, gb) * m.G + max(ga, gb)];
                if (pass == 0) {
                    if (convex_pair(m, ta, tb, ga, gb)) convex_convex(m, loc, 0, st, b, ga, gb, sink);
                } else if (ta == GT_PLANE && tb == GT_BOX) {
                    plane_box(m, loc, 0, ga, gb, sink);
                }
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
    if (placed > m.KMAX) loc.errbits[0] = loc.errbits[0] | RS_ERR_OVERFLOW_COLLISION_PAIRS;
    loc.n_contacts[0] = min(placed, m.KMAX);
    wave_sync();
}

template <bool kLds>
__global__ void __launch_bounds__(512) k_fused_wave(DModel m, DState st, FusedArgs fa) {
    extern __shared__ float lds_env[];
    const int w = (int)(threadIdx.x >> 6), lane = lane_id();
    // consecutive environments on one XCD (workgroups are dealt round-robin over the 8 XCDs), so that the planes'
    // cache lines, shared by neighbouring environments, stay in one L2
    const int nblk = (int)gridDim.x;
    int blk = (int)blockIdx.x;
    if ((nblk & 7) == 0) blk = (blk & 7) * (nblk >> 3) + (blk >> 3);
    const int b = blk * fa.epb + w;
    if (b >= st.B) return;  // whole wavefronts only
    float* wk = st.wk + (size_t)b * fa.work_stride;
    float* base = kLds ? lds_env + (size_t)w * fa.loc_floats : wk + fa.off_loc;
    const DState loc = make_local<kLds>(st, base, fa);
    const WorkRows R{wk + fa.off_jac, wk + fa.off_rowf, wk + fa.off_tmp, m.CMAX, m.D};

    stage_local<true>(m, st, loc, b, lane);
    wave_sync();
    // kernel_step_1: CRBA, L^T D L, actuation + passive + bias forces, acc_smooth (uniform on every lane)
    dev_mass_matrix(m, loc, 0);
    dev_factor_mass(m, loc, 0);
    dev_torque_passive(m, loc, 0);
    dev_update_acc(m, loc, 0, false);
    dev_rne_and_qacc(m, loc, 0);
    // Collider.detection
    if (m.enable_collision) {
        for (int g = lane; g < m.G; g += kWave) dev_aabb(m, loc, 0, g);
        wave_sync();
        if (m.P > 0) {
            const int n_old = loc.n_contacts[0];
            for (int i = lane; i < n_old; i += kWave) clear_contacts(st, b, i, i + 1);
            broadphase_sap(m, loc, 0, st, b);
            wave_sync();
            wave_narrowphase(m, loc, st, b, wk + fa.off_pc, lane);
        }
    }
    // ConstraintSolver.resolve
    if (!m.disable_constraint) {
        wave_constraints(m, loc, st, b, R, lane);
        wave_solve(m, loc, st, b, R, lane);
    }
    // kernel_step_2: cacc, semi-implicit Euler, forward kinematics and velocities
    dev_update_acc(m, loc, 0, true);
    dev_integrate(m, loc, 0);
    dev_kinematics(m, loc, 0, false);
    wave_sync();
    stage_local<false>(m, st, loc, b, lane);
}

}  // namespace rs
