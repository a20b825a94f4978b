// gt_inputs.hip -- the producers of the transport step's inputs, run by
// integrate_3d before the tracers (m3d.F90:446-548): start_macro (macro-step
// elevations, depths and mean transports), coordinates (layer heights on T,
// U and V points and the layer centres zc) and ww_momentum_3d (grid-related
// vertical velocity from layer continuity).  Cartesian metrics, kmin = 1.
//
// All of them are per-column recurrences or 2-D maps, HBM-bound: one thread
// per column, lanes along i so each level is one coalesced row per
// wavefront; the loads of a few levels are issued before their arithmetic.
// Every expression keeps the reference's order (built with -ffp-contract=off),
// so the results equal the CPU build's bit for bit.
#include "gt_kernels.h"

namespace gt {

namespace {

constexpr int kBx = 64, kBy = 4;
constexpr int kPre = 4;  // levels loaded ahead of their arithmetic

__device__ __forceinline__ double dmax2(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin2(double a, double b) { return a < b ? a : b; }

dim3 grid_cols(Dims d) {
  return dim3((unsigned)((d.NX + kBx - 1) / kBx), (unsigned)((d.NY + kBy - 1) / kBy));
}

// start_macro.F90:63-86: the U- and V-point elevations from the old (ssen)
// and new (z) T-point elevations, and the mean transports.  Runs before
// k_start_macro_t, which overwrites ssen.
__global__ void __launch_bounds__(kBx* kBy)
    k_start_macro_uv(Dims d, double split, double min_depth, const double* __restrict__ HU,
                     const double* __restrict__ HV, const double* __restrict__ z,
                     const double* __restrict__ ssen, double* __restrict__ ssuo,
                     double* __restrict__ ssun, double* __restrict__ ssvo,
                     double* __restrict__ ssvn, double* __restrict__ Dun,
                     double* __restrict__ Dvn, double* __restrict__ Uint,
                     double* __restrict__ Vint) {
  const int ii = blockIdx.x * kBx + threadIdx.x;
  const int jj = blockIdx.y * kBy + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long c = ii + (long)d.NX * jj;
  if (ii <= d.NX - 2) {
    // sseo(new) = ssen(old), ssen(new) = z
    double s = 0.25 * (ssen[c] + ssen[c + 1] + z[c] + z[c + 1]);
    ssuo[c] = ssun[c];
    s = dmax2(s, -HU[c] + min_depth);
    ssun[c] = s;
    Dun[c] = s + HU[c];
  }
  if (jj <= d.NY - 2) {
    double s = 0.25 * (ssen[c] + ssen[c + d.NX] + z[c] + z[c + d.NX]);
    ssvo[c] = ssvn[c];
    s = dmax2(s, -HV[c] + min_depth);
    ssvn[c] = s;
    Dvn[c] = s + HV[c];
  }
  Uint[c] = split * Uint[c];
  Vint[c] = split * Vint[c];
}

// start_macro.F90:54-61
__global__ void k_start_macro_t(long n, const double* __restrict__ H,
                                const double* __restrict__ z, double* __restrict__ sseo,
                                double* __restrict__ ssen, double* __restrict__ Dn) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (c >= n) return;
  sseo[c] = ssen[c];
  const double s = z[c];
  ssen[c] = s;
  Dn[c] = s + H[c];
}

// coordinates.F90:103-107 for one column, from its layer heights
__device__ __forceinline__ void zc_column(long c, long NXY, int kmax, double H,
                                          const double* __restrict__ hn, double* __restrict__ zc) {
  zc[c] = -H;
  double zp = -H + 0.5 * hn[c + NXY];
  double hp = hn[c + NXY];
  zc[c + NXY] = zp;
  for (int k = 2; k <= kmax; k++) {
    const double h = hn[c + NXY * (long)k];
    zp = zp + 0.5 * (hp + h);
    zc[c + NXY * (long)k] = zp;
    hp = h;
  }
}

// sigma_coordinates.F90:84-147.  EQUIV: h = D/kmax on every level k = 0..kmax
// (the whole-column assignment); otherwise h = D*dga(k) for k = 1..kmax.
template <bool EQUIV>
__global__ void __launch_bounds__(kBx* kBy)
    k_sigma(Dims d, double kmaxm1, const double* __restrict__ dga, const double* __restrict__ H,
            const double* __restrict__ HU, const double* __restrict__ HV,
            const double* __restrict__ sseo, const double* __restrict__ ssen,
            const double* __restrict__ ssuo, const double* __restrict__ ssun,
            const double* __restrict__ ssvo, const double* __restrict__ ssvn,
            double* __restrict__ ho, double* __restrict__ hn, double* __restrict__ huo,
            double* __restrict__ hun, double* __restrict__ hvo, double* __restrict__ hvn,
            double* __restrict__ zc) {
  const int ii = blockIdx.x * kBx + threadIdx.x;
  const int jj = blockIdx.y * kBy + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long c = ii + (long)d.NX * jj;
  const long NXY = d.NXY;
  const bool u = ii <= d.NX - 2, v = jj <= d.NY - 2;
  const double Do = sseo[c] + H[c], Dn = ssen[c] + H[c];
  const double Duo = u ? ssuo[c] + HU[c] : 0.0, Dun = u ? ssun[c] + HU[c] : 0.0;
  const double Dvo = v ? ssvo[c] + HV[c] : 0.0, Dvn = v ? ssvn[c] + HV[c] : 0.0;
  for (int k = EQUIV ? 0 : 1; k <= d.kmax; k++) {
    const double f = EQUIV ? kmaxm1 : dga[k];
    const long q = c + NXY * (long)k;
    ho[q] = Do * f;
    hn[q] = Dn * f;
    if (u) {
      huo[q] = Duo * f;
      hun[q] = Dun * f;
    }
    if (v) {
      hvo[q] = Dvo * f;
      hvn[q] = Dvn * f;
    }
  }
  if (zc) zc_column(c, NXY, d.kmax, H[c], hn, zc);
}

// general_coordinates.F90:82-102: the gamma levels of every column; a level
// below its predecessor sets *bad (the reference stops).  lev = sig | be.
__global__ void __launch_bounds__(kBx* kBy)
    k_gvc_gamma(Dims d, int kk, double d_gamma, double min_depth, const double* __restrict__ lev,
                const double* __restrict__ H, const double* __restrict__ sseo,
                double* __restrict__ gga, double* __restrict__ hn, int* bad) {
  const int ii = blockIdx.x * kBx + threadIdx.x;
  const int jj = blockIdx.y * kBy + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long c = ii + (long)d.NX * jj;
  const long NXY = d.NXY;
  const int K1 = d.kmax + 1;
  const double* sig = lev;
  const double* be = lev + K1;
  const double HH = dmax2(sseo[c] + H[c], min_depth);
  const double alpha =
      dmin2(((be[kk] - be[kk - 1]) - d_gamma / HH * (sig[kk] - sig[kk - 1])) /
                ((be[kk] - be[kk - 1]) - (sig[kk] - sig[kk - 1])),
            1.0);
  double gp = -1.0;
  gga[c] = gp;
  int folded = 0;
  for (int k = 1; k <= d.kmax; k++) {
    const double g = alpha * sig[k] + (1.0 - alpha) * be[k];
    gga[c + NXY * (long)k] = g;
    folded |= g < gp;
    // the initial layer heights, :109-116
    hn[c + NXY * (long)k] = HH * (g - gp);
    gp = g;
  }
  if (folded) atomicOr(bad, 1);
}

// general_coordinates.F90:119-139: the initial U and V layer heights
__global__ void __launch_bounds__(kBx* kBy)
    k_gvc_init_uv(Dims d, double min_depth, const double* __restrict__ gga,
                  const double* __restrict__ HU, const double* __restrict__ HV,
                  const double* __restrict__ ssuo, const double* __restrict__ ssvo,
                  double* __restrict__ huo, double* __restrict__ hun, double* __restrict__ hvo,
                  double* __restrict__ hvn) {
  const int ii = blockIdx.x * kBx + threadIdx.x;
  const int jj = blockIdx.y * kBy + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long c = ii + (long)d.NX * jj;
  const long NXY = d.NXY;
  const bool u = ii <= d.NX - 2, v = jj <= d.NY - 2;
  const double HHu = u ? dmax2(ssuo[c] + HU[c], min_depth) : 0.0;
  const double HHv = v ? dmax2(ssvo[c] + HV[c], min_depth) : 0.0;
  for (int k = 1; k <= d.kmax; k++) {
    const long q = c + NXY * (long)k, p = q - NXY;
    if (u) {
      const double h = HHu * 0.5 * (gga[q] - gga[p] + gga[q + 1] - gga[p + 1]);
      huo[q] = h;
      hun[q] = h;
    }
    if (v) {
      const double h = HHv * 0.5 * (gga[q] - gga[p] + gga[q + d.NX] - gga[p + d.NX]);
      hvo[q] = h;
      hvn[q] = h;
    }
  }
}

// general_coordinates.F90:146-210: relaxation of the T, U and V layer
// heights towards HH*dgamma; the top layer closes the column on the surface
// elevation.  Then zc (coordinates.F90:103-107).
__global__ void __launch_bounds__(kBx* kBy)
    k_gvc_step(Dims d, double relax_dt, double maxdepth, double d_gamma,
               const double* __restrict__ gga, const double* __restrict__ H,
               const double* __restrict__ HU, const double* __restrict__ HV,
               const double* __restrict__ ssen, const double* __restrict__ ssun,
               const double* __restrict__ ssvn, double* __restrict__ ho,
               double* __restrict__ hn, double* __restrict__ huo, double* __restrict__ hun,
               double* __restrict__ hvo, double* __restrict__ hvn, double* __restrict__ zc) {
  const int ii = blockIdx.x * kBx + threadIdx.x;
  const int jj = blockIdx.y * kBy + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long c = ii + (long)d.NX * jj;
  const long NXY = d.NXY;
  const long NX = d.NX;
  const int kmax = d.kmax;
  const bool u = ii <= d.NX - 2, v = jj <= d.NY - 2;
  const double Hc = H[c];
  // r = cord_relax/dt*H/maxdepth; relax_dt = cord_relax/dt
  const double rt = relax_dt * Hc / maxdepth;
  const double HH = ssen[c] + Hc;
  const bool uniform = HH < d_gamma;
  double ru = 0.0, HHu = 0.0, zu = 0.0, rv = 0.0, HHv = 0.0, zv = 0.0;
  if (u) {
    ru = relax_dt * HU[c] / maxdepth;
    zu = -HU[c];
    HHu = ssun[c] + HU[c];
  }
  if (v) {
    rv = relax_dt * HV[c] / maxdepth;
    zv = -HV[c];
    HHv = ssvn[c] + HV[c];
  }
  double zt = -Hc;
  double gc = gga[c], ge = u ? gga[c + 1] : 0.0, gn = v ? gga[c + NX] : 0.0;
  // zc from the new layer heights (kept when requested)
  double zcp = -Hc, hprev = 0.0;
  if (zc) zc[c] = -Hc;
  for (int k = 1; k <= kmax; k++) {
    const long q = c + NXY * (long)k;
    const bool top = k == kmax;
    const double gq = gga[q];
    double h;
    if (uniform) {
      h = HH / kmax;
      ho[q] = hn[q];
    } else {
      const double hold = hn[q];
      ho[q] = hold;
      if (!top) {
        h = (hold * rt + HH * (gq - gc)) / (rt + 1.0);
        zt = zt + h;
      } else {
        h = ssen[c] - zt;
      }
    }
    hn[q] = h;
    if (zc) {
      zcp = k == 1 ? -Hc + 0.5 * h : zcp + 0.5 * (hprev + h);
      zc[q] = zcp;
      hprev = h;
    }
    if (u) {
      const double geq = gga[q + 1];
      const double hold = hun[q];
      huo[q] = hold;
      double hu;
      if (!top) {
        hu = (hold * ru + HHu * 0.5 * (gq - gc + geq - ge)) / (ru + 1.0);
        zu = zu + hu;
      } else {
        hu = ssun[c] - zu;
      }
      hun[q] = hu;
      ge = geq;
    }
    if (v) {
      const double gnq = gga[q + NX];
      const double hold = hvn[q];
      hvo[q] = hold;
      double hv;
      if (!top) {
        hv = (hold * rv + HHv * 0.5 * (gq - gc + gnq - gn)) / (rv + 1.0);
        zv = zv + hv;
      } else {
        hv = ssvn[c] - zv;
      }
      hvn[q] = hv;
      gn = gnq;
    }
    gc = gq;
  }
}

// z-level "coordinates" leave the heights alone; only zc is formed
__global__ void __launch_bounds__(kBx* kBy)
    k_zc(Dims d, const double* __restrict__ H, const double* __restrict__ hn, double* zc) {
  const int ii = blockIdx.x * kBx + threadIdx.x;
  const int jj = blockIdx.y * kBy + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long c = ii + (long)d.NX * jj;
  zc_column(c, d.NXY, d.kmax, H[c], hn, zc);
}

// ww_momentum_3d.F90:66-107: upward from ww(0) on interior water columns
__global__ void __launch_bounds__(kBx* kBy)
    k_ww_momentum(Dims d, double dtm1, double dx, double dy, double ard1,
                  const int32_t* __restrict__ az, const double* __restrict__ uu,
                  const double* __restrict__ vv, const double* __restrict__ ho,
                  const double* __restrict__ hn, double* __restrict__ ww) {
  const int ii = blockIdx.x * kBx + threadIdx.x + 2;
  const int jj = blockIdx.y * kBy + threadIdx.y + 2;
  if (ii > d.NX - 3 || jj > d.NY - 3) return;
  const long c = ii + (long)d.NX * jj;
  if (az[c] != 1) return;
  const long NXY = d.NXY;
  const int n = d.kmax - 1;  // levels 1..kmax-1
  double w = ww[c];
  for (int k0 = 1; k0 <= n; k0 += kPre) {
    double a[kPre], b[kPre], e[kPre], f[kPre], g[kPre], h[kPre];
#pragma unroll
    for (int t = 0; t < kPre; t++) {
      if (k0 + t <= n) {
        const long q = c + NXY * (long)(k0 + t);
        a[t] = uu[q];
        b[t] = uu[q - 1];
        e[t] = vv[q];
        f[t] = vv[q - d.NX];
        g[t] = hn[q];
        h[t] = ho[q];
      }
    }
#pragma unroll
    for (int t = 0; t < kPre; t++) {
      if (k0 + t <= n) {
        w = w - (g[t] - h[t]) * dtm1 - (a[t] * dy - b[t] * dy + e[t] * dx - f[t] * dx) * ard1;
        ww[c + NXY * (long)(k0 + t)] = w;
      }
    }
  }
}

}  // namespace

hipError_t launch_start_macro(hipStream_t st, Dims d, double split, double min_depth,
                              const double* H, const double* HU, const double* HV,
                              const double* z, double* sseo, double* ssen, double* ssuo,
                              double* ssun, double* ssvo, double* ssvn, double* Dn, double* Dun,
                              double* Dvn, double* Uint, double* Vint) {
  hipLaunchKernelGGL(k_start_macro_uv, grid_cols(d), dim3(kBx, kBy), 0, st, d, split, min_depth,
                     HU, HV, z, ssen, ssuo, ssun, ssvo, ssvn, Dun, Dvn, Uint, Vint);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_start_macro_t, dim3((unsigned)((d.NXY + 255) / 256)), dim3(256), 0, st,
                     d.NXY, H, z, sseo, ssen, Dn);
  return hipGetLastError();
}

hipError_t launch_sigma(hipStream_t st, Dims d, bool equiv, double kmaxm1, const double* dga,
                        const double* H, const double* HU, const double* HV, const double* sseo,
                        const double* ssen, const double* ssuo, const double* ssun,
                        const double* ssvo, const double* ssvn, double* ho, double* hn,
                        double* huo, double* hun, double* hvo, double* hvn, double* zc) {
  if (equiv)
    hipLaunchKernelGGL(k_sigma<true>, grid_cols(d), dim3(kBx, kBy), 0, st, d, kmaxm1, dga, H, HU,
                       HV, sseo, ssen, ssuo, ssun, ssvo, ssvn, ho, hn, huo, hun, hvo, hvn, zc);
  else
    hipLaunchKernelGGL(k_sigma<false>, grid_cols(d), dim3(kBx, kBy), 0, st, d, kmaxm1, dga, H,
                       HU, HV, sseo, ssen, ssuo, ssun, ssvo, ssvn, ho, hn, huo, hun, hvo, hvn,
                       zc);
  return hipGetLastError();
}

hipError_t launch_gvc_init(hipStream_t st, Dims d, int kk, double d_gamma, double min_depth,
                           const double* lev, const double* H, const double* HU,
                           const double* HV, const double* sseo, const double* ssuo,
                           const double* ssvo, double* gga, double* hn, double* huo,
                           double* hun, double* hvo, double* hvn, int* bad) {
  hipLaunchKernelGGL(k_gvc_gamma, grid_cols(d), dim3(kBx, kBy), 0, st, d, kk, d_gamma, min_depth,
                     lev, H, sseo, gga, hn, bad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_gvc_init_uv, grid_cols(d), dim3(kBx, kBy), 0, st, d, min_depth, gga, HU,
                     HV, ssuo, ssvo, huo, hun, hvo, hvn);
  return hipGetLastError();
}

hipError_t launch_gvc_step(hipStream_t st, Dims d, double relax_dt, double maxdepth,
                           double d_gamma, const double* gga, const double* H, const double* HU,
                           const double* HV, const double* ssen, const double* ssun,
                           const double* ssvn, double* ho, double* hn, double* huo, double* hun,
                           double* hvo, double* hvn, double* zc) {
  hipLaunchKernelGGL(k_gvc_step, grid_cols(d), dim3(kBx, kBy), 0, st, d, relax_dt, maxdepth,
                     d_gamma, gga, H, HU, HV, ssen, ssun, ssvn, ho, hn, huo, hun, hvo, hvn, zc);
  return hipGetLastError();
}

hipError_t launch_zc(hipStream_t st, Dims d, const double* H, const double* hn, double* zc) {
  hipLaunchKernelGGL(k_zc, grid_cols(d), dim3(kBx, kBy), 0, st, d, H, hn, zc);
  return hipGetLastError();
}

hipError_t launch_ww_momentum(hipStream_t st, Dims d, double dt, double dx, double dy,
                              double ard1, const int32_t* az, const double* uu, const double* vv,
                              const double* ho, const double* hn, double* ww) {
  if (d.NX < 5 || d.NY < 5 || d.kmax < 2) return hipSuccess;
  const dim3 g((unsigned)((d.NX - 4 + kBx - 1) / kBx), (unsigned)((d.NY - 4 + kBy - 1) / kBy));
  hipLaunchKernelGGL(k_ww_momentum, g, dim3(kBx, kBy), 0, st, d, 1.0 / dt, dx, dy, ard1, az, uu,
                     vv, ho, hn, ww);
  return hipGetLastError();
}

}  // namespace gt
