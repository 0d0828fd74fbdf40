// Host-side launchers of the fused substep kernels. Each kernel class is instantiated in a translation unit of its own
// (rs_fused_t16.hip, rs_fused_t24.hip, rs_fused_gen.hip, rs_fused_multi.hip) so that the library builds in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "rs_fused.h"

namespace rs {

// k_fused_wave for models with at most 16 dofs (row tile of 16), occupancy 0, 3 or 4
void launch_fused_tile16(bool use_lds, int occupancy, dim3 grid, dim3 block, size_t lds_bytes, hipStream_t s,
                         const DModel& m, const DState& st, const FusedArgs& fa);
// k_fused_wave for models with 17 to 24 dofs (row tile of 24), occupancy 0 or 3
void launch_fused_tile24(bool use_lds, int occupancy, dim3 grid, dim3 block, size_t lds_bytes, hipStream_t s,
                         const DModel& m, const DState& st, const FusedArgs& fa);
// k_fused_wave with the generic solver (more than 24 dofs)
void launch_fused_generic(bool use_lds, dim3 grid, dim3 block, size_t lds_bytes, hipStream_t s, const DModel& m,
                          const DState& st, const FusedArgs& fa);
// k_fused_multi: one fused substep of several (model, slab) segments
void launch_fused_multi(bool use_lds, dim3 grid, dim3 block, size_t lds_bytes, hipStream_t s,
                        const FusedSegment* segs, int n_segs);

}  // namespace rs
