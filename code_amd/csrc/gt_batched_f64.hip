// gt_batched_f64.hip -- the FP64 (reference-exact) instantiations of the
// fused HALFSPLIT kernel (gt_fused_impl.h).
#include "gt_fused_impl.h"

namespace gt {

void fused_tile_shape(int tile, int* TX, int* TY) {
  switch (tile) {
    case 1: *TX = 64; *TY = 4; break;
    case 2: *TX = 16; *TY = 8; break;
    default: *TX = 32; *TY = 8; break;
  }
}

bool fused_supported(int hscheme, int vscheme) {
  return hscheme == vscheme &&
         (hscheme == UPSTREAM || hscheme == P2 || hscheme == SUPERBEE || hscheme == MUSCL ||
          hscheme == P2_PDM);
}

hipError_t launch_fused_f64(hipStream_t st, const FusedArgs<double>& a, int tile, bool adv) {
  if (adv) return dispatch_scheme<double, 1, true, true, 32, 8, 512>(st, a);
  switch (tile) {
    case 1: return dispatch_scheme<double, 1, true, false, 64, 4, 512>(st, a);
    case 2: return dispatch_scheme<double, 1, true, false, 16, 8, 256>(st, a);
    default: return dispatch_scheme<double, 1, true, false, 32, 8, 512>(st, a);
  }
}

}  // namespace gt

namespace gt {
hipError_t launch_fused_f32_g1(hipStream_t st, const FusedArgs<float>& a, int tile);
hipError_t launch_fused_f32_g3(hipStream_t st, const FusedArgs<float>& a, int tile);
hipError_t launch_fused_f32_g6(hipStream_t st, const FusedArgs<float>& a, int tile);

hipError_t launch_fused_f32(hipStream_t st, const FusedArgs<float>& a, int tile, int group) {
  switch (group) {
    case 1: return launch_fused_f32_g1(st, a, tile);
    case 6: return launch_fused_f32_g6(st, a, tile);
    default: return launch_fused_f32_g3(st, a, tile);
  }
}
}  // namespace gt
