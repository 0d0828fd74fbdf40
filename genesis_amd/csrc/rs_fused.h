// Fused substep: every stage of one environment's substep in a single kernel launch (one lane per environment),
// so intermediates stay in the caches between stages and a substep costs one launch instead of ten.
#pragma once
#include "rs_collision.h"
#include "rs_constraint.h"
#include "rs_dyn.h"

namespace rs {

constexpr int kFusedBlock = 64;

__global__ void __launch_bounds__(kFusedBlock) k_fused_substep(DModel m, DState st) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= st.B) return;
    dev_mass_matrix(m, st, b);
    dev_factor_mass(m, st, b);
    dev_torque_passive(m, st, b);
    dev_update_acc(m, st, b, false);
    dev_rne_and_qacc(m, st, b);
    if (m.enable_collision) {
        for (int g = 0; g < m.G; g++) dev_aabb(m, st, b, g);
        if (m.P > 0) {
            dev_broadphase(m, st, b);
            dev_narrowphase(m, st, b);
        }
    }
    if (!m.disable_constraint) {
        dev_equality_constraints(m, st, b);
        dev_inequality_constraints(m, st, b);
        dev_solve(m, st, b);
    }
    dev_update_acc(m, st, b, true);
    dev_integrate(m, st, b);
    dev_kinematics(m, st, b, false);
}

inline int fused_substep(const DModel& m, const DState& st, hipStream_t s) {
    hipLaunchKernelGGL(k_fused_substep, dim3((st.B + kFusedBlock - 1) / kFusedBlock), dim3(kFusedBlock), 0, s, m, st);
    return hipGetLastError() == hipSuccess ? RS_OK : RS_E_DEVICE;
}

}  // namespace rs
