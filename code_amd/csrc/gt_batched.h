// gt_batched.h -- the HALFSPLIT transport step as two strip kernels
// (gt_strip_impl.h), used by the single-tracer FP64 path and the batched
// multi-tracer FP32/FP64 entry points.
#pragma once

#include "gt_device.h"

namespace gt {

// columns a strip writes: a wavefront covers 64 columns, the outer two on each
// side are the ghost columns of the two consecutive i-direction sub-steps
constexpr int kStripOut = 60;

// packed per-cell mask bits (one byte per cell, per mask set H/U/V)
enum : uint8_t { MB_UFLUX = 1, MB_VFLUX = 2, MB_UUPD = 4, MB_VUPD = 8, MB_WCOL = 16 };
// in kernel registers only: the lane's column is one the strip writes
enum : int { MB_XOUT = 32 };

// one domain of a launch.  A launch covers the work items of one or more
// domains (the uneven subdomains of a batched run) back to back.  A work item
// is one strip (60 columns x `rows` rows) at one level for one tracer group.
template <typename TF>
struct FusedDomain {
  Dims d;
  int nsx;            // strips along i
  int nsy[2];         // strips along j, of the first and the second kernel
  long item_begin[2]; // first launch-wide work item of this domain, per kernel
  const TF* fin;      // tracer t of this domain starts at fin + t*tstride
  TF* fout;           // output, same layout (may equal fin)
  long tstride;
  TF* f2;             // work: tracers after u/2 + v/2, same layout as fin
  double* d2;         // work: layer heights after u/2 + v/2 (one field)
  double* a2;         // work (ADV): advective term accumulated over u/2 + v/2
  const double *uu, *vv, *ww, *hu, *hv, *ho;
  const uint8_t* flags;  // MB_* bits of the mask set of the grid
  double* hires;      // optional: end of the layer-height trajectory
  double* advres;     // optional (FP64, one tracer): accumulated advective term
  // optional: the budget partials of the second kernel, one (hi, lo) pair
  // per output q (tracer q, the volume at q = ntracers) and strip-level sl:
  // bpart[2*(q*nsl + sl) + {0,1}], nsl = nsx*nsy[1]*kmax
  double* bpart;
  double dx, dy, ard1;
};

constexpr int kMaxFusedDomains = 8;

// passed by value as the kernel argument (< 4 KB)
template <typename TF>
struct FusedArgs {
  FusedDomain<TF> doms[kMaxFusedDomains];
  int ndom;
  long nitems[2];  // work items over all domains, per kernel
  int ngroups;     // tracer groups
  int ntracers;
  int rows[2];     // rows per strip of the first (u/2 v/2) and second (w v/2 u/2) kernel
  double dt, AH;
  int hscheme, vscheme, kshift;
};

// rows per strip for a tuning index: 0 -> 16, 1 -> 32, 2 -> 8 for both
// kernels; 3..11 -> the first kernel {16, 32, 8}[(t-3)%3], the second
// {16, 32, 8}[(t-3)/3] (multiples of 4: the row march is unrolled four times)
void strip_rows(int tile, int rows[2]);
// the strip kernels are instantiated for hscheme == vscheme in UPSTREAM..P2_PDM
bool fused_supported(int hscheme, int vscheme);
// group: tracers per wavefront.  FP64: 1 or 2 (adv needs 1); FP32: 1, 3, 6 or 9.
// The caller fills everything but ngroups and the item ranges, which
// strip_plan sets.
template <typename TF>
void strip_plan(FusedArgs<TF>& a, int group);
hipError_t launch_fused_f64(hipStream_t st, const FusedArgs<double>& a, int group, bool adv);
hipError_t launch_fused_f32(hipStream_t st, const FusedArgs<float>& a, int group);

}  // namespace gt
