// gt_batched_f64.hip -- the FP64 (reference-exact) instantiations of the
// HALFSPLIT strip kernels (gt_strip_impl.h), and the launch plan.
#include "gt_strip_impl.h"

namespace gt {

void strip_rows(int tile, int rows[2]) {
  static const int r[3] = {16, 32, 8};
  if (tile >= 3 && tile < 12) {
    rows[0] = r[(tile - 3) % 3];
    rows[1] = r[(tile - 3) / 3];
  } else {
    rows[0] = rows[1] = r[(tile >= 0 && tile < 3) ? tile : 0];
  }
}

bool fused_supported(int hscheme, int vscheme) {
  return hscheme == vscheme &&
         (hscheme == UPSTREAM || hscheme == P2 || hscheme == SUPERBEE || hscheme == MUSCL ||
          hscheme == P2_PDM);
}

template void strip_plan<double>(FusedArgs<double>&, int);
template void strip_plan<float>(FusedArgs<float>&, int);

hipError_t launch_fused_f64(hipStream_t st, const FusedArgs<double>& a, int group, bool adv) {
  if (adv) return dispatch_scheme<double, 1, true, true>(st, a);
  if (group == 2) return dispatch_scheme<double, 2, true, false>(st, a);
  return dispatch_scheme<double, 1, true, false>(st, a);
}

}  // namespace gt

namespace gt {
hipError_t launch_fused_f32_g1(hipStream_t st, const FusedArgs<float>& a);
hipError_t launch_fused_f32_g3(hipStream_t st, const FusedArgs<float>& a);
hipError_t launch_fused_f32_g6(hipStream_t st, const FusedArgs<float>& a);
hipError_t launch_fused_f32_g9(hipStream_t st, const FusedArgs<float>& a);

hipError_t launch_fused_f32(hipStream_t st, const FusedArgs<float>& a, int group) {
  switch (group) {
    case 1: return launch_fused_f32_g1(st, a);
    case 6: return launch_fused_f32_g6(st, a);
    case 9: return launch_fused_f32_g9(st, a);
    default: return launch_fused_f32_g3(st, a);
  }
}
}  // namespace gt
