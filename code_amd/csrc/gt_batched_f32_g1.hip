// gt_batched_f32_g1.hip -- FP32 instantiations of the HALFSPLIT strip
// kernels (gt_strip_impl.h) with 1 tracer(s) per wavefront.
#include "gt_strip_impl.h"

namespace gt {

hipError_t launch_fused_f32_g1(hipStream_t st, const FusedArgs<float>& a) {
  return dispatch_scheme<float, 1, false, false>(st, a);
}

}  // namespace gt
