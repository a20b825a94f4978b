// gt_batched_f32_g1.hip -- FP32 instantiations of the fused HALFSPLIT kernel
// (gt_fused_impl.h) with 1 tracer(s) per workgroup.
#include "gt_fused_impl.h"

namespace gt {

hipError_t launch_fused_f32_g1(hipStream_t st, const FusedArgs<float>& a, int tile) {
  switch (tile) {
    case 1: return dispatch_scheme<float, 1, false, false, 64, 4, 512>(st, a);
    case 2: return dispatch_scheme<float, 1, false, false, 16, 8, 256>(st, a);
    default: return dispatch_scheme<float, 1, false, false, 32, 8, 512>(st, a);
  }
}

}  // namespace gt
