// gt_adjoint.hip -- reverse mode of the Blumberg-Mellor internal pressure
// and the linear EOS (FP64); the advection's reverse mode is in
// gt_substeps.hip.
//
// There is no counterpart in getm-model/code; the forward operators are
// ip_blumberg_mellor (src/3d/ip_blumberg_mellor.F90) and the linear case of
// do_eqstate (src/3d/eqstate.F90:127-149,252-261).  Both are linear in the
// buoyancy (layer heights held fixed); their transposes are gathered per
// cell or column, so no atomics are needed.
#include "gt_adjoint.h"

namespace gt {

// Adjoint of ip_blumberg_mellor (ip_blumberg_mellor.F90:99-149) with respect
// to buoy (zz, hun, hvn fixed), gathered per cell: cell (i,j) receives from
// the u-points i-1 (as the i+1 side) and i (as the i side), and the v-points
// j-1 and j.  Each u-point's downward recurrence prgr_k = prgr_{k+1} + ...
// is transposed by an upward sweep of its adjoint.
__device__ void bm_point_adj(const double* buoy_unused, const double* zz, const double* hh,
                             const double* ibar, long p, long step, long NXY, int kmax,
                             double dm1, double side, double* out_k /* length kmax+1 */) {
  (void)buoy_unused;
  // pbar_k = hh_k ibar_k + pbar_{k-1}   (prgr_k feeds prgr_{k-1})
  double pbar_prev = 0.0;  // pbar_{k-1}
  for (int k = 1; k <= kmax; k++) {
    double hk = hh[p + NXY * k];
    double pbar = hk * ibar[p + NXY * k] + pbar_prev;
    // terms where level k's grd and buoyl appear:
    //  prgr_k (k<kmax): +0.25 (hun_k + hun_{k+1}) grd_k  and  +dxz_k buoyl_k
    //  prgr_kmax:       +0.5 hun_kmax grd_kmax
    //  prgr_{k-1} (k>=2): +0.25 (hun_{k-1}+hun_k) grd_k  and  -dxz_{k-1} buoyl_k
    double gbar = 0.0, lbar = 0.0;
    if (k < kmax) {
      gbar += pbar * 0.5 * 0.5 * (hk + hh[p + NXY * (k + 1)]);
      double dz = (zz[p + step + NXY * k] - zz[p + NXY * k]) * dm1;
      lbar += pbar * dz;
    } else {
      gbar += pbar * 0.5 * hk;
    }
    if (k >= 2) {
      gbar += pbar_prev * 0.5 * 0.5 * (hh[p + NXY * (k - 1)] + hk);
      double dz = (zz[p + step + NXY * (k - 1)] - zz[p + NXY * (k - 1)]) * dm1;
      lbar -= pbar_prev * dz;
    }
    // grd = (b(side+) - b(side0)) dm1,  buoyl = 0.5 (b(+) + b(0))
    out_k[k] += side * gbar * dm1 + 0.5 * lbar;
    pbar_prev = pbar;
  }
}

__global__ void k_bm_adj(Dims d, const int32_t* __restrict__ au, const int32_t* __restrict__ av,
                         const double* __restrict__ zz, const double* __restrict__ hun,
                         const double* __restrict__ hvn, const double* __restrict__ ixbar,
                         const double* __restrict__ iybar, double* bbar, double dxm1,
                         double dym1) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const long NX = d.NX;
  const long c0 = ii + NX * jj;
  const int kmax = d.kmax;
  double acc[64 + 1];
  for (int k = 0; k <= kmax; k++) acc[k] = 0.0;
  auto interior = [&](int i, int j) { return i >= 2 && i <= d.NX - 3 && j >= 2 && j <= d.NY - 3; };
  // u-point (ii, jj): this cell is the i side (sign -1 in grd)
  if (interior(ii, jj) && au[c0] >= 1)
    bm_point_adj(nullptr, zz, hun, ixbar, c0, 1, NXY, kmax, dxm1, -1.0, acc);
  // u-point (ii-1, jj): this cell is the i+1 side
  if (ii >= 1 && interior(ii - 1, jj) && au[c0 - 1] >= 1)
    bm_point_adj(nullptr, zz, hun, ixbar, c0 - 1, 1, NXY, kmax, dxm1, 1.0, acc);
  if (interior(ii, jj) && av[c0] >= 1)
    bm_point_adj(nullptr, zz, hvn, iybar, c0, NX, NXY, kmax, dym1, -1.0, acc);
  if (jj >= 1 && interior(ii, jj - 1) && av[c0 - NX] >= 1)
    bm_point_adj(nullptr, zz, hvn, iybar, c0 - NX, NX, NXY, kmax, dym1, 1.0, acc);
  for (int k = 0; k <= kmax; k++) bbar[c0 + NXY * k] = acc[k];
}

// linear EOS transpose: T_bar = x dtr0 buoy_bar, S_bar = x dsr0 buoy_bar
__global__ void k_eos_linear_adj(long n, long off, const double* __restrict__ bbar, double* Tbar,
                                 double* Sbar, double dtr0, double dsr0) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  long c = off + q;
  const double g = (double)9.81f, rho_0 = (double)1025.f;
  const double x = -g / rho_0;
  if (Tbar) Tbar[c] = x * dtr0 * bbar[c];
  if (Sbar) Sbar[c] = x * dsr0 * bbar[c];
}

static inline dim3 g2(const Dims& d, int nz) { return dim3((d.NX + 63) / 64, (d.NY + 3) / 4, nz); }

hipError_t launch_bm_adj(hipStream_t st, Dims d, const int32_t* au, const int32_t* av,
                         const double* zz, const double* hun, const double* hvn,
                         const double* ixbar, const double* iybar, double* bbar, double dx,
                         double dy) {
  if (d.kmax > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bm_adj, g2(d, 1), dim3(64, 4), 0, st, d, au, av, zz, hun, hvn, ixbar,
                     iybar, bbar, 1.0 / dx, 1.0 / dy);
  return hipGetLastError();
}

hipError_t launch_eos_linear_adj(hipStream_t st, Dims d, const double* bbar, double* Tbar,
                                 double* Sbar, double dtr0, double dsr0) {
  long n = d.NXY * (long)d.kmax;
  hipLaunchKernelGGL(k_eos_linear_adj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     d.NXY, bbar, Tbar, Sbar, dtr0, dsr0);
  return hipGetLastError();
}

}  // namespace gt
