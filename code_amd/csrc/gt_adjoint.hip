// gt_adjoint.hip -- reverse mode of the split advection step and of the
// Blumberg-Mellor internal pressure with the linear EOS (FP64).
//
// There is no counterpart in getm-model/code; the forward operators are
// adv_split_u/v/w (src/2d/adv_split_u.F90, adv_split_v.F90,
// src/3d/adv_split_w.F90), ip_blumberg_mellor (src/3d/ip_blumberg_mellor.F90)
// and the linear case of do_eqstate (src/3d/eqstate.F90:127-149,252-261).
// Velocities and layer heights are held fixed, so every sub-step is a
// (piecewise) linear map of f:
//   f_new(i) = (Dio(i) f(i) - dt advn(i)) / Dn(i),
//   advn(i)  = s (F(i) - F(i-1)),  s = splitfac*dy*ard1,
//   F(X)     = U(X) rec(f(X-1..X+2)) - AH DU (f(X+1) - f(X)) / dx.
// Its transpose is evaluated as a gather: every thread owns one cell and
// sums the contributions of the (up to) four faces whose stencil holds it,
// so no atomics are needed.  The limiter's derivative is the derivative of
// the branch the forward evaluation took (one-sided at kinks).
#include "gt_adjoint.h"

namespace gt {

// d rec / d (fuu, fu, fd) at the forward point (advection.F90:785-879)
struct RecGrad {
  double duu, du, dd;
};

__device__ __forceinline__ RecGrad reconstruct_grad(int scheme, double cfl, double fuu, double fu,
                                                    double fd) {
  const double one6th = 1.0 / 6;
  const double g = 0.5 * (1.0 - cfl);
  double deltaf = fd - fu, deltafu = fu - fuu;
  if (scheme == P2) {  // linear in (fuu, fu, fd) on both branches
    double x = one6th * (1.0 - 2.0 * cfl);
    return {-g * (0.5 - x), 1.0 - g * (0.5 + x) + g * (0.5 - x), g * (0.5 + x)};
  }
  if (scheme == UPSTREAM || !(deltaf * deltafu > 0.0)) return {0.0, 1.0, 0.0};
  double r = deltafu / deltaf;
  double phi, dphi;
  switch (scheme) {
    case P2_PDM: {
      double x = one6th * (1.0 - 2.0 * cfl);
      double a = 2.0 * r / (cfl + 1.e-10), b = (0.5 + x) + (0.5 - x) * r, c = 2.0 / (1.0 - cfl);
      // the forward takes min(min(a, b), c)
      double m = a < b ? a : b;
      double dm = a < b ? 2.0 / (cfl + 1.e-10) : (0.5 - x);
      if (c < m) { phi = c; dphi = 0.0; } else { phi = m; dphi = dm; }
      break;
    }
    case SUPERBEE: {
      double a = 2.0 * r < 1.0 ? 2.0 * r : 1.0, da = 2.0 * r < 1.0 ? 2.0 : 0.0;
      double b = r < 2.0 ? r : 2.0, db = r < 2.0 ? 1.0 : 0.0;
      if (a > b) { phi = a; dphi = da; } else { phi = b; dphi = db; }
      break;
    }
    case MUSCL: {
      double a = 2.0 * r, b = 0.5 * (1.0 + r);
      double m = a < b ? a : b, dm = a < b ? 2.0 : 0.5;
      if (2.0 < m) { phi = 2.0; dphi = 0.0; } else { phi = m; dphi = dm; }
      break;
    }
    default:
      return {0.0, 1.0, 0.0};
  }
  // rec = fu + g phi(r) deltaf,  r = (fu-fuu)/deltaf
  return {-g * dphi, 1.0 + g * (dphi * (1.0 + r) - phi), g * (phi - dphi * r)};
}

// partials of the face flux at face X (cells X-1..X+2 along the sweep) with
// respect to the four stencil cells, written to out[0..3] (cells X-1..X+2)
__device__ __forceinline__ void face_flux_grad(int scheme, const double* f, long c, long st,
                                               const uint8_t* mflux, long m, long mst,
                                               double vel, double Dvel, double dti, double dlen,
                                               double AH, double out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (!mflux[m]) return;
  if (vel > 0.0) {
    bool lim = scheme != UPSTREAM && mflux[m - mst];
    if (lim) {
      double cfl = vel / Dvel * dti / dlen;
      RecGrad gr = reconstruct_grad(scheme, cfl, f[c - st], f[c], f[c + st]);
      out[0] = vel * gr.duu;
      out[1] = vel * gr.du;
      out[2] = vel * gr.dd;
    } else {
      out[1] = vel;
    }
  } else {
    bool lim = scheme != UPSTREAM && mflux[m + mst];
    if (lim) {
      double cfl = -vel / Dvel * dti / dlen;
      RecGrad gr = reconstruct_grad(scheme, cfl, f[c + 2 * st], f[c + st], f[c]);
      out[3] = vel * gr.duu;
      out[2] = vel * gr.du;
      out[1] = vel * gr.dd;
    } else {
      out[2] = vel;
    }
  }
  if (AH > 0.0) {
    double k = AH * Dvel / dlen;
    out[1] += k;
    out[2] -= k;
  }
}

// transpose of one horizontal sub-step.  dir 0: u (stride 1, faces by i),
// dir 1: v (stride NX).  f: the state before the sub-step; Dio/Dn: layer
// heights before/after; lam_in: adjoint of the output; lam_out: of the input
__global__ void k_split_uv_adj(Dims d, int k0, int nslice, int dir, const double* __restrict__ f,
                               const double* __restrict__ lam_in, double* __restrict__ lam_out,
                               const double* __restrict__ Dio, const double* __restrict__ Dn,
                               const double* __restrict__ vel, const double* __restrict__ Dvel,
                               const uint8_t* __restrict__ mflux,
                               const uint8_t* __restrict__ mupd, double dt, double splitfac,
                               int scheme, double AH, double dlen_flux, double dlen_face,
                               double ard1) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  int s = blockIdx.z;
  if (ii >= d.NX || jj >= d.NY || s >= nslice) return;
  const long o = d.NXY * (long)(k0 + s);
  const long c2 = ii + (long)d.NX * jj;
  const long st = dir == 0 ? 1 : d.NX;
  const int pos = dir == 0 ? ii : jj;           // coordinate along the sweep
  const int n = dir == 0 ? d.NX : d.NY;
  const int sub = scheme == UPSTREAM ? 0 : 1;
  const double dti = splitfac * dt;
  auto updated = [&](int p, long q2) { return p >= 1 + sub && p <= n - 2 - sub && mupd[q2]; };
  // mu(q): adjoint of F entering cell q's update with + sign
  auto mu = [&](int p, long q2) -> double {
    if (p < 0 || p >= n || !updated(p, q2)) return 0.0;
    long q = o + q2;
    return lam_in[q] * (-dt / Dn[q]) * (splitfac * dlen_face * ard1);
  };
  double acc;
  if (updated(pos, c2)) acc = lam_in[o + c2] * Dio[o + c2] / Dn[o + c2];
  else acc = lam_in[o + c2];
  // faces X = pos-2 .. pos+1 hold cell pos at stencil slot pos-X+1
  for (int X = pos - 2; X <= pos + 1; X++) {
    if (X < sub || X > n - 2 - sub) continue;
    long x2 = c2 + (long)(X - pos) * st;
    double Fbar = mu(X, x2) - mu(X + 1, x2 + st);
    if (Fbar == 0.0) continue;
    double gr[4];
    face_flux_grad(scheme, f + o, x2, st, mflux, x2, st, vel[o + x2], Dvel[o + x2], dti,
                   dlen_flux, AH, gr);
    acc += Fbar * gr[pos - X + 1];
  }
  lam_out[o + c2] = acc;
}

// transpose of the vertical sub-step (itersmax = 1), one thread per column
__global__ void k_split_w_adj(Dims d, const double* __restrict__ f,
                              const double* __restrict__ lam_in, double* __restrict__ lam_out,
                              const double* __restrict__ hio, const double* __restrict__ hn,
                              const double* __restrict__ ww, const int32_t* __restrict__ az,
                              double dt, double splitfac, int scheme, int kshift) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const int kmax = d.kmax;
  const long c0 = ii + (long)d.NX * jj;
  const bool col = ii <= d.NX - 2 && jj <= d.NY - 2 && az[c0] == 1 && scheme != NOADV;
  if (!col) {
    for (int k = 0; k <= kmax; k++) lam_out[c0 + NXY * k] = lam_in[c0 + NXY * k];
    return;
  }
  const double dti = splitfac * dt;
#define AT(a, k) (a)[c0 + NXY * (long)(k)]
  auto updated = [&](int k) { return k >= 1 && k <= kmax - kshift; };
  auto mu = [&](int k) -> double {
    if (k < 0 || k > kmax || !updated(k)) return 0.0;
    return AT(lam_in, k) * (-dt / AT(hn, k)) * splitfac;
  };
  // gradient of wflux(K) w.r.t. f at K-1..K+2
  auto wgrad = [&](int K, double out[4]) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    double w = AT(ww, K);
    if (w > 0.0) {
      bool lim = scheme != UPSTREAM && K > 1 - kshift;
      if (lim) {
        double cfl = fabs(w) * dti / (0.5 * (AT(hio, K) + AT(hio, K + 1)));
        RecGrad g = reconstruct_grad(scheme, cfl, AT(f, K - 1), AT(f, K), AT(f, K + 1));
        out[0] = w * g.duu; out[1] = w * g.du; out[2] = w * g.dd;
      } else {
        out[1] = w;
      }
    } else {
      bool lim = scheme != UPSTREAM && K < kmax - 1;
      if (lim) {
        double cfl = fabs(w) * dti / (0.5 * (AT(hio, K) + AT(hio, K + 1)));
        RecGrad g = reconstruct_grad(scheme, cfl, AT(f, K + 2), AT(f, K + 1), AT(f, K));
        out[3] = w * g.duu; out[2] = w * g.du; out[1] = w * g.dd;
      } else {
        out[2] = w;
      }
    }
  };
  for (int k = 0; k <= kmax; k++) {
    double acc = updated(k) ? AT(lam_in, k) * AT(hio, k) / AT(hn, k) : AT(lam_in, k);
    for (int K = k - 2; K <= k + 1; K++) {
      if (K < 1 - kshift || K > kmax - 1) continue;
      double Wbar = mu(K) - mu(K + 1);
      if (Wbar == 0.0) continue;
      double g[4];
      wgrad(K, g);
      acc += Wbar * g[k - K + 1];
    }
    AT(lam_out, k) = acc;
  }
#undef AT
}

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

hipError_t launch_split_uv_adj(hipStream_t st, Dims d, int k0, int nslice, int dir,
                               const double* f, const double* lam_in, double* lam_out,
                               const double* Dio, const double* Dn, const double* vel,
                               const double* Dvel, const uint8_t* mflux, const uint8_t* mupd,
                               double dt, double splitfac, int scheme, double AH,
                               double dlen_flux, double dlen_face, double ard1) {
  hipLaunchKernelGGL(k_split_uv_adj, g2(d, nslice), dim3(64, 4), 0, st, d, k0, nslice, dir, f,
                     lam_in, lam_out, Dio, Dn, vel, Dvel, mflux, mupd, dt, splitfac, scheme, AH,
                     dlen_flux, dlen_face, ard1);
  return hipGetLastError();
}

hipError_t launch_split_w_adj(hipStream_t st, Dims d, const double* f, const double* lam_in,
                              double* lam_out, const double* hio, const double* hn,
                              const double* ww, const int32_t* az, double dt, double splitfac,
                              int scheme, int kshift) {
  hipLaunchKernelGGL(k_split_w_adj, g2(d, 1), dim3(64, 4), 0, st, d, f, lam_in, lam_out, hio, hn,
                     ww, az, dt, splitfac, scheme, kshift);
  return hipGetLastError();
}

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
