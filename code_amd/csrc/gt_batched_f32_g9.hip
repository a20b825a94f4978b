// gt_batched_f32_g9.hip -- FP32 instantiations of the HALFSPLIT strip
// kernels (gt_strip_impl.h) with 9 tracer(s) per wavefront.
#include "gt_strip_impl.h"

namespace gt {

hipError_t launch_fused_f32_g9(hipStream_t st, const FusedArgs<float>& a) {
  return dispatch_scheme<float, 9, false, false>(st, a);
}

}  // namespace gt
