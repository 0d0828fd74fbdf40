// Fused substep kernels of models with at most 16 dofs (row tile of 16): private region in LDS or in the work field,
// occupancy 0 (compiler's register budget), 3 or 4 wavefronts per SIMD.
#include "rs_launch.h"

namespace rs {

void launch_fused_tile16(bool use_lds, int occ, dim3 grid, dim3 block, size_t lds, hipStream_t s, const DModel& m,
                         const DState& st, const FusedArgs& fa) {
    if (use_lds) {
        if (occ == 4) hipLaunchKernelGGL((k_fused_wave<true, 4, 1>), grid, block, lds, s, m, st, fa);
        else if (occ == 3) hipLaunchKernelGGL((k_fused_wave<true, 3, 1>), grid, block, lds, s, m, st, fa);
        else hipLaunchKernelGGL((k_fused_wave<true, 0, 1>), grid, block, lds, s, m, st, fa);
    } else {
        if (occ == 4) hipLaunchKernelGGL((k_fused_wave<false, 4, 1>), grid, block, lds, s, m, st, fa);
        else if (occ == 3) hipLaunchKernelGGL((k_fused_wave<false, 3, 1>), grid, block, lds, s, m, st, fa);
        else hipLaunchKernelGGL((k_fused_wave<false, 0, 1>), grid, block, lds, s, m, st, fa);
    }
}

}  // namespace rs
