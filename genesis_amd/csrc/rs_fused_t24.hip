// Fused substep kernels of models with 17 to 24 dofs (row tile of 24): occupancy 0 or 3.
#include "rs_launch.h"

namespace rs {

void launch_fused_tile24(bool use_lds, int occ, dim3 grid, dim3 block, size_t lds, hipStream_t s, const DModel& m,
                         const DState& st, const FusedArgs& fa) {
    if (use_lds) {
        if (occ == 3) hipLaunchKernelGGL((k_fused_wave<true, 3, 4>), grid, block, lds, s, m, st, fa);
        else hipLaunchKernelGGL((k_fused_wave<true, 0, 4>), grid, block, lds, s, m, st, fa);
    } else {
        if (occ == 3) hipLaunchKernelGGL((k_fused_wave<false, 3, 4>), grid, block, lds, s, m, st, fa);
        else hipLaunchKernelGGL((k_fused_wave<false, 0, 4>), grid, block, lds, s, m, st, fa);
    }
}

}  // namespace rs
