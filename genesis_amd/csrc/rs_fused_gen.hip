// Fused substep kernels of models with more than 24 dofs (generic solver).
#include "rs_launch.h"

namespace rs {

void launch_fused_generic(bool use_lds, dim3 grid, dim3 block, size_t lds, hipStream_t s, const DModel& m,
                          const DState& st, const FusedArgs& fa) {
    if (use_lds) hipLaunchKernelGGL((k_fused_wave<true, 0, 0>), grid, block, lds, s, m, st, fa);
    else hipLaunchKernelGGL((k_fused_wave<false, 0, 0>), grid, block, lds, s, m, st, fa);
}

}  // namespace rs
