// gt_batched.h -- the fused HALFSPLIT transport kernel (gt_batched.hip).
#pragma once

#include "gt_device.h"

namespace gt {

// one domain of a fused launch.  A launch covers the tile tables of one or
// more domains (the uneven subdomains of a batched run) back to back.
template <typename TF>
struct FusedDomain {
  Dims d;
  int ntx, nty;       // tiles along i and j
  long tile_begin;    // first launch-wide tile index of this domain
  const TF* fin;   // tracer t of this domain starts at fin + t*tstride
  TF* fout;        // output, same layout; must not overlap fin
  long tstride;
  const double *uu, *vv, *ww, *hu, *hv, *ho;
  const uint8_t *m_uflux, *m_vflux, *m_uupd, *m_vupd;
  const int32_t* az;
  double* hires;   // optional: end of the layer-height trajectory
  double* advres;  // optional (FP64, one tracer): accumulated advective term
  double dx, dy, ard1;
};

constexpr int kMaxFusedDomains = 8;

// passed by value as the kernel argument (< 4 KB)
template <typename TF>
struct FusedArgs {
  FusedDomain<TF> doms[kMaxFusedDomains];
  int ndom;
  long ntiles;     // tiles over all domains
  int ngroups;     // tracer groups (set by the launcher)
  int ntracers;
  double dt, AH;
  int hscheme, vscheme, kshift;
};

// tile shapes: 0 -> 32x8, 1 -> 64x4, 2 -> 16x8
void fused_tile_shape(int tile, int* TX, int* TY);
// the fused kernel is instantiated for hscheme == vscheme in UPSTREAM..P2_PDM
bool fused_supported(int hscheme, int vscheme);
hipError_t launch_fused_f64(hipStream_t st, const FusedArgs<double>& a, int tile, bool adv);
// group: tracers per workgroup (1, 3 or 6)
hipError_t launch_fused_f32(hipStream_t st, const FusedArgs<float>& a, int tile, int group);

}  // namespace gt
