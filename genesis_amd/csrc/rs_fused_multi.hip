// Shared fused launches of several (model, slab) segments (serving).
#include "rs_launch.h"

namespace rs {

void launch_fused_multi(bool use_lds, dim3 grid, dim3 block, size_t lds, hipStream_t s, const FusedSegment* segs,
                        int n_segs) {
    if (use_lds) hipLaunchKernelGGL(k_fused_multi<true>, grid, block, lds, s, segs, n_segs);
    else hipLaunchKernelGGL(k_fused_multi<false>, grid, block, lds, s, segs, n_segs);
}

}  // namespace rs
