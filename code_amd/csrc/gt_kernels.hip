// gt_kernels.hip -- per-sub-step FP64 kernels of the transport step.
//
// One kernel per reference routine, each a faithful restatement of its loop
// bounds and expression order (citations per kernel, getm-model/code).  The
// reference updates fi in place while it reads neighbours of f; when f and
// fi alias, the host passes a second buffer as fi_dst and swaps afterwards,
// and every cell the reference does not touch is copied through.
#include "gt_ip.h"
#include "gt_kernels.h"

namespace gt {

// ---------------------------------------------------------------------------
// a2: adv_split_u, src/2d/adv_split_u.F90:205-286.  Grid: (ii, jj, slice).
// The interface flux at i is a pure function of f, U, DU and the masks, so
// each thread recomputes the two fluxes its update needs instead of storing
// the uflux array.
__device__ __forceinline__ double flux_u(const double* f, const double* U, const double* DU,
                                         const uint8_t* mflux, long c, int scheme, double dti,
                                         double dx, double AH) {
  if (!mflux[c]) return 0.0;
  double fu, fuu = 0.0, fd = 0.0, cfl = 0.0;
  bool use_limiter = false;
  double u = U[c];
  if (u > 0.0) {
    fu = f[c];
    if (scheme != UPSTREAM) use_limiter = mflux[c - 1];
    if (use_limiter) {
      cfl = u / DU[c] * dti / dx;
      fuu = f[c - 1];
      fd = f[c + 1];
    }
  } else {
    fu = f[c + 1];
    if (scheme != UPSTREAM) use_limiter = mflux[c + 1];
    if (use_limiter) {
      cfl = -u / DU[c] * dti / dx;
      fuu = f[c + 2];
      fd = f[c];
    }
  }
  if (use_limiter) fu = reconstruct<double>(scheme, cfl, fuu, fu, fd);
  double fl = u * fu;
  if (AH > 0.0) fl = fl - AH * DU[c] * (f[c + 1] - f[c]) / dx;
  return fl;
}

__global__ void k_split_u(Dims d, int nslice, const double* __restrict__ f,
                          const double* fi_src, double* fi_dst, double* Di, double* adv,
                          const double* __restrict__ U, const double* __restrict__ DU,
                          const uint8_t* __restrict__ mflux, const uint8_t* __restrict__ mupd,
                          double dt, double splitfac, int scheme, double AH, double dx,
                          double dy, double ard1) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  int s = blockIdx.z;
  if (ii >= d.NX || jj >= d.NY || s >= nslice) return;
  long o = (long)s * d.NXY;
  long c2 = ii + (long)d.NX * jj;
  long c = o + c2;
  const int isub = (scheme == UPSTREAM) ? 0 : 1;
  bool upd = ii >= 1 + isub && ii <= d.NX - 2 - isub && mupd[c2];
  if (upd) {
    const double dti = splitfac * dt;
    double fc = flux_u(f + o, U + o, DU + o, mflux, c2, scheme, dti, dx, AH);
    double fw = flux_u(f + o, U + o, DU + o, mflux, c2 - 1, scheme, dti, dx, AH);
    double Dio = Di[c];
    double Dn = Dio - dti * (U[c] * dy - U[c - 1] * dy) * ard1;
    Di[c] = Dn;
    double advn = splitfac * (fc * dy - fw * dy) * ard1;
    fi_dst[c] = (Dio * fi_src[c] - dt * advn) / Dn;
    adv[c] = adv[c] + advn;
  } else if (fi_dst != fi_src) {
    fi_dst[c] = fi_src[c];
  }
}

// a3: adv_split_v, src/2d/adv_split_v.F90:63-136
__device__ __forceinline__ double flux_v(const double* f, const double* V, const double* DV,
                                         const uint8_t* mflux, long c, long NX, int scheme,
                                         double dti, double dy, double AH) {
  if (!mflux[c]) return 0.0;
  double fu, fuu = 0.0, fd = 0.0, cfl = 0.0;
  bool use_limiter = false;
  double v = V[c];
  if (v > 0.0) {
    fu = f[c];
    if (scheme != UPSTREAM) use_limiter = mflux[c - NX];
    if (use_limiter) {
      cfl = v / DV[c] * dti / dy;
      fuu = f[c - NX];
      fd = f[c + NX];
    }
  } else {
    fu = f[c + NX];
    if (scheme != UPSTREAM) use_limiter = mflux[c + NX];
    if (use_limiter) {
      cfl = -v / DV[c] * dti / dy;
      fuu = f[c + 2 * NX];
      fd = f[c];
    }
  }
  if (use_limiter) fu = reconstruct<double>(scheme, cfl, fuu, fu, fd);
  double fl = v * fu;
  if (AH > 0.0) fl = fl - AH * DV[c] * (f[c + NX] - f[c]) / dy;
  return fl;
}

__global__ void k_split_v(Dims d, int nslice, const double* __restrict__ f,
                          const double* fi_src, double* fi_dst, double* Di, double* adv,
                          const double* __restrict__ V, const double* __restrict__ DV,
                          const uint8_t* __restrict__ mflux, const uint8_t* __restrict__ mupd,
                          double dt, double splitfac, int scheme, double AH, double dx,
                          double dy, double ard1) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  int s = blockIdx.z;
  if (ii >= d.NX || jj >= d.NY || s >= nslice) return;
  long o = (long)s * d.NXY;
  long c2 = ii + (long)d.NX * jj;
  long c = o + c2;
  const int jsub = (scheme == UPSTREAM) ? 0 : 1;
  bool upd = jj >= 1 + jsub && jj <= d.NY - 2 - jsub && mupd[c2];
  if (upd) {
    const double dti = splitfac * dt;
    const long NX = d.NX;
    double fc = flux_v(f + o, V + o, DV + o, mflux, c2, NX, scheme, dti, dy, AH);
    double fs = flux_v(f + o, V + o, DV + o, mflux, c2 - NX, NX, scheme, dti, dy, AH);
    double Dio = Di[c];
    double Dn = Dio - dti * (V[c] * dx - V[c - NX] * dx) * ard1;
    Di[c] = Dn;
    double advn = splitfac * (fc * dx - fs * dx) * ard1;
    fi_dst[c] = (Dio * fi_src[c] - dt * advn) / Dn;
    adv[c] = adv[c] + advn;
  } else if (fi_dst != fi_src) {
    fi_dst[c] = fi_src[c];
  }
}

// ---------------------------------------------------------------------------
// a4: adv_split_w, src/3d/adv_split_w.F90:85-353, itersmax <= 1.
// One thread per water column; the flux at level k only needs the original
// f at k-1..k+2 and the original hi at k, k+1, and level k is updated right
// after its top flux is known, so one upward sweep does the whole column.
// f is never written (fi_dst != f is guaranteed by the host).
__global__ void k_split_w(Dims d, const double* __restrict__ f, const double* fi_src,
                          double* fi_dst, double* hi, double* adv,
                          const double* __restrict__ ww, const int32_t* __restrict__ az,
                          double dt, double splitfac, int scheme, int kshift) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const int kmax = d.kmax;
  long c0 = ii + (long)d.NX * jj;
  bool col = ii <= d.NX - 2 && jj <= d.NY - 2 && az[c0] == 1;
  const bool copy = fi_dst != fi_src;
  if (!col) {
    if (copy)
      for (int k = 0; k <= kmax; k++) fi_dst[c0 + NXY * k] = fi_src[c0 + NXY * k];
    return;
  }
  if (copy) {
    fi_dst[c0] = fi_src[c0];
    if (kshift) fi_dst[c0 + NXY * kmax] = fi_src[c0 + NXY * kmax];
  }
  const double dti = splitfac * dt;
#define AT(a, k) (a)[c0 + NXY * (long)(k)]
  auto wflux_at = [&](int k) -> double {
    double w = AT(ww, k);
    double fu, fuu = 0.0, fd = 0.0;
    bool use_limiter = false;
    if (w > 0.0) {
      fu = AT(f, k);
      if (scheme != UPSTREAM) use_limiter = k > 1 - kshift;
      if (use_limiter) {
        fuu = AT(f, k - 1);
        fd = AT(f, k + 1);
      }
    } else {
      fu = AT(f, k + 1);
      if (scheme != UPSTREAM) use_limiter = k < kmax - 1;
      if (use_limiter) {
        fuu = AT(f, k + 2);
        fd = AT(f, k);
      }
    }
    if (use_limiter) {
      double cfl = fabs(w) * dti / (0.5 * (AT(hi, k) + AT(hi, k + 1)));
      fu = reconstruct<double>(scheme, cfl * 1.0, fuu, fu, fd);
    }
    return w * fu;
  };
  double wprev = kshift ? wflux_at(0) : 0.0;
  double wwprev = AT(ww, 0);
  for (int k = 1; k <= kmax - kshift; k++) {
    double wk = (k <= kmax - 1) ? wflux_at(k) : 0.0;
    double wwk = AT(ww, k);
    double hio = AT(hi, k);
    double hn = hio - dti * (wwk - wwprev);
    AT(hi, k) = hn;
    double advn = splitfac * (wk - wprev);
    AT(fi_dst, k) = (hio * AT(fi_src, k) - dt * advn) / hn;
    AT(adv, k) = AT(adv, k) + advn;
    wprev = wk;
    wwprev = wwk;
  }
#undef AT
}

// a4 with sub-cycling (itersmax > 1): the reference's column algorithm
// literally, with its column work arrays in a scratch field laid out
// [slot][k][column] so that neighbouring threads stay coalesced.
// Scratch slots: 0 fo, 1 fiaux, 2 hiaux, 3 advaux, 4 cfl0, 5 cfls, 6 wflux.
__global__ void k_split_w_iter(Dims d, const double* __restrict__ f, const double* fi_src,
                               double* fi_dst, double* hi, double* adv,
                               const double* __restrict__ ww, const int32_t* __restrict__ az,
                               double* scratch, unsigned long long* nwarn, double dt,
                               double splitfac, int scheme, int kshift, int itersmax) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const int kmax = d.kmax;
  const long N3 = NXY * (kmax + 1);
  long c0 = ii + (long)d.NX * jj;
  bool col = ii <= d.NX - 2 && jj <= d.NY - 2 && az[c0] == 1;
  const bool copy = fi_dst != fi_src;
  if (!col) {
    if (copy)
      for (int k = 0; k <= kmax; k++) fi_dst[c0 + NXY * k] = fi_src[c0 + NXY * k];
    return;
  }
#define AT(a, k) (a)[c0 + NXY * (long)(k)]
#define SC(slot, k) scratch[(slot) * N3 + c0 + NXY * (long)(k)]
  const double dti = splitfac * dt;
  for (int k = 1 - kshift; k <= kmax - 1; k++)
    SC(4, k) = fabs(AT(ww, k)) * dti / (0.5 * (AT(hi, k) + AT(hi, k + 1)));
  int iters = 1;
  double itersm1 = 1.0, dtik = dti, splitfack = splitfac;
  for (int k = 0; k <= kmax; k++) SC(0, k) = AT(f, k);
  SC(6, 0) = 0.0;
  SC(6, kmax) = 0.0;
  int fa = 0;  // faux slot: 0 = fo, 1 = fiaux
  int it = 1;
  // every restart strictly raises iters (bounded by itersmax), so the loop
  // runs at most itersmax*(itersmax+1)/2 sweeps
  while (it <= iters) {
    if (it == 1) {
      for (int k = 0; k <= kmax; k++) {
        SC(1, k) = AT(fi_src, k);
        SC(2, k) = AT(hi, k);
        SC(3, k) = AT(adv, k);
      }
      fa = 0;
    }
    if (it == 1) {
      for (int k = 0; k <= kmax; k++) SC(5, k) = SC(4, k);
    } else if (scheme != UPSTREAM || iters < itersmax) {
      for (int k = 1 - kshift; k <= kmax - 1; k++)
        SC(5, k) = fabs(AT(ww, k)) * dti / (0.5 * (SC(2, k) + SC(2, k + 1)));
    }
    if (iters < itersmax) {
      int mx = INT_MIN;
      for (int k = 1 - kshift; k <= kmax - 1; k++) {
        double cc = ceil(SC(5, k));
        int ci = cc > 2147483647.0 ? 2147483647 : (int)cc;
        mx = ci > mx ? ci : mx;
      }
      int iters_new = mx > 1 ? mx : 1;
      if (iters_new > iters) {
        if (iters_new > itersmax) {
          atomicAdd(nwarn, 1ull);
          iters = itersmax;
        } else {
          iters = iters_new;
        }
        itersm1 = 1.0 / iters;
        dtik = dti * itersm1;
        splitfack = splitfac * itersm1;
        if (it > 1) {
          it = 1;
          continue;
        }
      }
    }
    // fluxes from faux (slot 6 holds wflux(0:kmax), ends preset to 0),
    // then the update of the aux column in place
    for (int k = 1 - kshift; k <= kmax - 1; k++) {
      double w = AT(ww, k);
      double fu, fuu = 0.0, fd = 0.0;
      bool use_limiter = false;
      if (w > 0.0) {
        fu = SC(fa, k);
        if (scheme != UPSTREAM) use_limiter = k > 1 - kshift;
        if (use_limiter) {
          fuu = SC(fa, k - 1);
          fd = SC(fa, k + 1);
        }
      } else {
        fu = SC(fa, k + 1);
        if (scheme != UPSTREAM) use_limiter = k < kmax - 1;
        if (use_limiter) {
          fuu = SC(fa, k + 2);
          fd = SC(fa, k);
        }
      }
      if (use_limiter) fu = reconstruct<double>(scheme, SC(5, k) * itersm1, fuu, fu, fd);
      SC(6, k) = w * fu;
    }
    for (int k = 1; k <= kmax - kshift; k++) {
      double hio = SC(2, k);
      SC(2, k) = hio - dtik * (AT(ww, k) - AT(ww, k - 1));
      double advn = splitfack * (SC(6, k) - SC(6, k - 1));
      SC(1, k) = (hio * SC(1, k) - dt * advn) / SC(2, k);
      SC(3, k) = SC(3, k) + advn;
    }
    fa = 1;
    it++;
  }
  if (copy) fi_dst[c0] = fi_src[c0];
  for (int k = 1; k <= kmax - kshift; k++) {
    AT(fi_dst, k) = SC(1, k);
    AT(hi, k) = SC(2, k);
    AT(adv, k) = SC(3, k);
  }
  if (kshift) AT(fi_dst, kmax) = AT(fi_src, kmax);
#undef SC
#undef AT
}

// NOSPLIT finalisation, advection_3d.F90:416-423 (and advection.F90:450-458
// per slice): f = (ho*f - dt*adv)/hi on mask_finalise, k = 1..kmax.
__global__ void k_finalise(Dims d, int k0, int nk, double* f, const double* __restrict__ ho,
                           const double* __restrict__ adv, const double* __restrict__ hi,
                           const uint8_t* __restrict__ mfin, double dt) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  int s = blockIdx.z;
  if (ii >= d.NX || jj >= d.NY || s >= nk) return;
  long c2 = ii + (long)d.NX * jj;
  if (!mfin[c2]) return;
  long c = c2 + d.NXY * (long)(k0 + s);
  f[c] = (ho[c] * f[c] - dt * adv[c]) / hi[c];
}

// ---------------------------------------------------------------------------
// a7: adv_upstream_2dh, src/2d/adv_upstream_2dh.F90:70-160.  Interior only.
__device__ __forceinline__ double uflux_2dh(const double* f, const double* U, const double* Dn,
                                            const int32_t* az, long c, long c2, double AH,
                                            double dx) {
  double u = U[c];
  double fl = u > 0.0 ? u * f[c] : u * f[c + 1];
  if (AH > 0.0 && az[c2] > 0 && az[c2 + 1] > 0)
    fl = fl - AH * (f[c + 1] - f[c]) / dx * 0.5 * (Dn[c + 1] + Dn[c]);
  return fl;
}
__device__ __forceinline__ double vflux_2dh(const double* f, const double* V, const double* Dn,
                                            const int32_t* az, long c, long c2, long NX,
                                            double AH, double dy) {
  double v = V[c];
  double fl = v > 0.0 ? v * f[c] : v * f[c + NX];
  if (AH > 0.0 && az[c2] > 0 && az[c2 + NX] > 0)
    fl = fl - AH * (f[c + NX] - f[c]) / dy * 0.5 * (Dn[c + NX] + Dn[c]);
  return fl;
}

__global__ void k_upstream_2dh(Dims d, int nslice, const double* __restrict__ f,
                               const double* fi_src, double* fi_dst, double* Di, double* adv,
                               const double* __restrict__ U, const double* __restrict__ V,
                               const double* __restrict__ Dn, const int32_t* __restrict__ az,
                               double dt, double AH, double dx, double dy, double ard1) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  int s = blockIdx.z;
  if (ii >= d.NX || jj >= d.NY || s >= nslice) return;
  const long NX = d.NX;
  long o = (long)s * d.NXY;
  long c2 = ii + NX * jj;
  long c = o + c2;
  bool upd = ii >= 2 && ii <= d.NX - 3 && jj >= 2 && jj <= d.NY - 3 && az[c2] == 1;
  if (!upd) {
    if (fi_dst != fi_src) fi_dst[c] = fi_src[c];
    return;
  }
  const double* fs = f + o;
  double mn = 100000 * 1.0, mx = -100000 * 1.0;
  for (int q = -1; q <= 1; q++)
    if (az[c2 + q] == 1) {
      double v = fs[c2 + q];
      if (v > mx) mx = v;
      if (v < mn) mn = v;
    }
  for (int q = -1; q <= 1; q += 2)
    if (az[c2 + q * NX] == 1) {
      double v = fs[c2 + q * NX];
      if (v > mx) mx = v;
      if (v < mn) mn = v;
    }
  double ue = uflux_2dh(fs, U + o, Dn + o, az, c2, c2, AH, dx);
  double uw = uflux_2dh(fs, U + o, Dn + o, az, c2 - 1, c2 - 1, AH, dx);
  double vn = vflux_2dh(fs, V + o, Dn + o, az, c2, c2, NX, AH, dy);
  double vs = vflux_2dh(fs, V + o, Dn + o, az, c2 - NX, c2 - NX, NX, AH, dy);
  double Dio = Di[c];
  double Dnew = Dio - dt * (U[c] * dy - U[c - 1] * dy + V[c] * dx - V[c - NX] * dx) * ard1;
  Di[c] = Dnew;
  double advn = (ue * dy - uw * dy + vn * dx - vs * dx) * ard1;
  double fv = (Dio * fi_src[c] - dt * advn) / Dnew;
  adv[c] = adv[c] + advn;
  if (fv > mx) fv = mx;
  if (fv < mn) fv = mn;
  fi_dst[c] = fv;
}

// ---------------------------------------------------------------------------
// a12: linear EOS and buoyancy, src/3d/eqstate.F90:127-149,252-261; all
// cells (land sentinels included), k = 1..kmax.
__global__ void k_eos_linear(long n, long off, const double* __restrict__ T,
                             const double* __restrict__ S, double* rho, double* buoy, double T0,
                             double S0, double dtr0, double dsr0) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  long c = off + q;
  if (rho) rho[c] = (double)1025.f + dtr0 * (T[c] - T0) + dsr0 * (S[c] - S0);
  buoy[c] = eos_linear_buoy(T[c], S[c], T0, S0, dtr0, dsr0);
}

// levels whose loads a column march issues together: the per-column
// recurrences below stay sequential, but a block's operands arrive in one
// memory round trip instead of one per level
// (4 for Blumberg-Mellor, whose eight operands per level would otherwise
// need more registers than a wavefront beside the advection's can take)
constexpr int kColBlock = 8;
constexpr int kBmBlock = 4;

// ---------------------------------------------------------------------------
// a9: ip_blumberg_mellor, src/3d/ip_blumberg_mellor.F90:67-150.
// zz: interface heights over [imin..imax+1]x[jmin..jmax+1] where az>=1.
__global__ __launch_bounds__(256) void k_bm_zz(Dims d, const int32_t* __restrict__ az, const double* __restrict__ H,
                        const double* __restrict__ hn, double* __restrict__ zz) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  long c0 = ii + (long)d.NX * jj;
  zz[c0] = 0.0;
  if (ii < 2 || ii > d.NX - 2 || jj < 2 || jj > d.NY - 2 || az[c0] < 1) return;
  double z = -H[c0] + hn[c0 + d.NXY];
  zz[c0 + d.NXY] = z;
  for (int kb = 2; kb <= d.kmax; kb += kColBlock) {
    double q[kColBlock];
#pragma unroll
    for (int u = 0; u < kColBlock; u++) q[u] = hn[c0 + d.NXY * min(kb + u, d.kmax)];
    // levels past kmax leave z alone (a select, not a branch, so the loads
    // stay ahead of their uses)
#pragma unroll
    for (int u = 0; u < kColBlock; u++) {
      const bool on = kb + u <= d.kmax;
      const double zn = z + q[u];
      z = on ? zn : z;
      if (on) zz[c0 + d.NXY * (kb + u)] = z;
    }
  }
}

// idpdx and idpdy (:99-149) in one downward loop per column (the column's
// own buoy and zz are loaded once for both directions); each recurrence is
// the reference's, in its order.
__global__ __launch_bounds__(256) void k_bm_idpd(Dims d, const int32_t* __restrict__ au, const int32_t* __restrict__ av,
                          const double* __restrict__ buoy, const double* __restrict__ zz,
                          const double* __restrict__ hun, const double* __restrict__ hvn,
                          double* __restrict__ idpdx, double* __restrict__ idpdy, double dxm1,
                          double dym1) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const int kmax = d.kmax;
  long c0 = ii + (long)d.NX * jj;
  idpdx[c0] = 0.0;
  idpdy[c0] = 0.0;
  if (ii < 2 || ii > d.NX - 3 || jj < 2 || jj > d.NY - 3) return;
  const bool ux = au[c0] >= 1, vy = av[c0] >= 1;
  if (!ux && !vy) return;
#define AT(a, q, k) (a)[(q) + NXY * (long)(k)]
  const long e = c0 + 1, n = c0 + d.NX;
  double b0 = AT(buoy, c0, kmax), be = AT(buoy, e, kmax), bn = AT(buoy, n, kmax);
  double grdlx = (be - b0) * dxm1, buoylx = 0.5 * (be + b0), hkx = AT(hun, c0, kmax);
  double prgrx = grdlx * 0.5 * hkx;
  double grdly = (bn - b0) * dym1, buoyly = 0.5 * (bn + b0), hky = AT(hvn, c0, kmax);
  double prgry = grdly * 0.5 * hky;
  if (ux) AT(idpdx, c0, kmax) = hkx * prgrx;
  if (vy) AT(idpdy, c0, kmax) = hky * prgry;
  for (int kb = kmax - 1; kb >= 1; kb -= kBmBlock) {
    // the operands of kBmBlock levels are loaded first (one memory round
    // trip per block instead of per level); levels below 1 repeat level 1
    double qb0[kBmBlock], qbe[kBmBlock], qbn[kBmBlock], qz0[kBmBlock], qze[kBmBlock],
        qzn[kBmBlock], qhu[kBmBlock], qhv[kBmBlock];
#pragma unroll
    for (int u = 0; u < kBmBlock; u++) {
      const int k = max(kb - u, 1);
      qb0[u] = AT(buoy, c0, k);
      qbe[u] = AT(buoy, e, k);
      qbn[u] = AT(buoy, n, k);
      qz0[u] = AT(zz, c0, k);
      qze[u] = AT(zz, e, k);
      qzn[u] = AT(zz, n, k);
      qhu[u] = AT(hun, c0, k);
      qhv[u] = AT(hvn, c0, k);
    }
#pragma unroll
    for (int u = 0; u < kBmBlock; u++) {
      const int k = kb - u;
      if (k < 1) break;
      b0 = qb0[u];
      be = qbe[u];
      bn = qbn[u];
      const double z0 = qz0[u], ze = qze[u], zn = qzn[u];
      {
        double grdu = grdlx;
        grdlx = (be - b0) * dxm1;
        double buoyu = buoylx;
        buoylx = 0.5 * (be + b0);
        double dxz = (ze - z0) * dxm1;
        double hkk = qhu[u];
        prgrx = prgrx + 0.5 * (grdu + grdlx) * 0.5 * (hkk + hkx) - dxz * (buoyu - buoylx);
        if (ux) AT(idpdx, c0, k) = hkk * prgrx;
        hkx = hkk;
      }
      {
        double grdu = grdly;
        grdly = (bn - b0) * dym1;
        double buoyu = buoyly;
        buoyly = 0.5 * (bn + b0);
        double dyz = (zn - z0) * dym1;
        double hkk = qhv[u];
        prgry = prgry + 0.5 * (grdu + grdly) * 0.5 * (hkk + hky) - dyz * (buoyu - buoyly);
        if (vy) AT(idpdy, c0, k) = hkk * prgry;
        hky = hkk;
      }
    }
  }
#undef AT
}

// ---------------------------------------------------------------------------
// a10: ip_shchepetkin_mcwilliams, src/3d/ip_shchepetkin_mcwilliams.F90:71-288
// (constants and harmonic means in gt_ip.h)

// One thread per column: the layer centres zz over the full halo range where
// az>=1 (:90-99), then, on [imin..imax+1]x[jmin..jmax+1], the vertical
// harmonic means and the pressure integral P (:108-146), reading back the
// column's zz just written (one launch instead of two, zz and hn stay in
// cache between the two loops).
__global__ __launch_bounds__(256) void k_sm_column(Dims d, const int32_t* __restrict__ az, const double* __restrict__ H,
                            const double* __restrict__ hn, const double* __restrict__ buoy,
                            double* __restrict__ zz, double* __restrict__ P) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const int kmax = d.kmax;
  long c0 = ii + (long)d.NX * jj;
  zz[c0] = 0.0;
  if (az[c0] < 1) return;
#define AT(a, k) (a)[c0 + NXY * (long)(k)]
  double zt, zm;  // zz at kmax and kmax-1, kept for the second loop
  {
    double z = -H[c0] + 0.5 * AT(hn, 1);
    AT(zz, 1) = z;
    double hprev = AT(hn, 1);
    zm = z;
    for (int kb = 2; kb <= kmax; kb += kColBlock) {
      double q[kColBlock];
#pragma unroll
      for (int u = 0; u < kColBlock; u++) q[u] = AT(hn, min(kb + u, kmax));
#pragma unroll
      for (int u = 0; u < kColBlock; u++) {
        const bool on = kb + u <= kmax;  // a select, as in k_bm_zz
        const double hk = q[u];
        const double zn = z + 0.5 * (hprev + hk);
        zm = on ? z : zm;
        z = on ? zn : z;
        hprev = on ? hk : hprev;
        if (on) AT(zz, kb + u) = z;
      }
    }
    zt = z;
  }
  if (ii < 2 || ii > d.NX - 2 || jj < 2 || jj > d.NY - 2) return;
  // vertical differences dR(k) = b(k+1)-b(k), dZ likewise, for k in
  // 1..kmax-1, extended by dR(kmax) = dR(kmax-1) and dR(0) = dR(1) (:108-121);
  // the window b(k+1), b(k), zz(k+1), zz(k) moves down in registers
  auto hmZ = [&](double a, double b) -> double { return sm_hmZ(a, b); };
  double b1 = AT(buoy, kmax), b0 = AT(buoy, kmax - 1);
  double z1 = zt, z0 = zm;
  const double ht = AT(hn, kmax);
  double cff1 = 1.0 / (z1 - z0);
  double cff2 = 0.5 * (b1 - b0) * 0.5 * ht * cff1;
  double p = (b1 + cff2) * 0.5 * ht;
  AT(P, kmax) = p;
  double rk = b1 - b0, zk = z1 - z0;             // dR, dZ at kmax-1 (= at kmax)
  double hR1 = sm_hmR(rk, rk), hZ1 = hmZ(zk, zk);  // harmonic means at kmax
  for (int kb = kmax - 1; kb >= 1; kb -= kColBlock) {
    // buoy and zz at k-1 for the block's levels, loaded first (level 1
    // stands in below level 1, unused)
    double qb[kColBlock], qz[kColBlock];
#pragma unroll
    for (int u = 0; u < kColBlock; u++) {
      const int l = max(kb - u - 1, 1);
      qb[u] = AT(buoy, l);
      qz[u] = AT(zz, l);
    }
#pragma unroll
    for (int u = 0; u < kColBlock; u++) {
      const int k = kb - u;
      if (k < 1) break;
      double bm = 0.0, zmv = 0.0, rkm, zkm;
      if (k >= 2) {
        bm = qb[u];
        zmv = qz[u];
        rkm = b0 - bm;
        zkm = z0 - zmv;
      } else {
        rkm = b1 - b0;  // dR(0) = dR(1)
        zkm = z1 - z0;
      }
      double hR = sm_hmR(rk, rkm), hZ = hmZ(zk, zkm);
      p = p + 0.5 * ((b1 + b0) * (z1 - z0) -
                     SM_ONEFIFTH * ((hR1 - hR) * (z1 - z0 - SM_C12 * (hZ1 + hZ)) -
                                    (hZ1 - hZ) * (b1 - b0 - SM_C12 * (hR1 + hR))));
      AT(P, k) = p;
      hR1 = hR;
      hZ1 = hZ;
      rk = rkm;
      zk = zkm;
      b1 = b0;
      b0 = bm;
      z1 = z0;
      z0 = zmv;
    }
  }
#undef AT
}


// idpdx (:153-211) and idpdy (:214-287).  One wavefront per (strip of 60
// columns x 16 rows, level): lane l owns column x = 60*sx - 1 + l, the
// i-neighbours x-1..x+2 come from adjacent lanes (DPP), the j-neighbours
// y-1..y+2 from a four-row window of the march, so every value is loaded
// once.  Level 0 is zeroed; the arithmetic is the per-cell formula's.
constexpr int kIpRows = 16;
__global__ __launch_bounds__(256) void k_sm_idpd_strip(
    Dims d, int nsx, int nsy, const int32_t* __restrict__ au, const int32_t* __restrict__ av,
    const double* __restrict__ buoy, const double* __restrict__ zz, const double* __restrict__ P,
    const double* __restrict__ hun, const double* __restrict__ hvn,
    const double* __restrict__ sseo, double* __restrict__ idpdx, double* __restrict__ idpdy,
    double dxm1, double dym1) {
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long item = (long)blockIdx.x * 4 + wave;
  const long nitems = (long)nsx * nsy * (d.kmax + 1);
  if (item >= nitems) return;
  const int k = (int)(item % (d.kmax + 1));
  const int sx = (int)((item / (d.kmax + 1)) % nsx);
  const int sy = (int)(item / ((long)(d.kmax + 1) * nsx));
  const int lane = threadIdx.x & 63;
  const int NX = d.NX, NY = d.NY;
  const long NXY = d.NXY;
  const int x = 60 * sx - 1 + lane;
  const int y0 = sy * kIpRows;
  const bool xout = lane >= 1 && lane <= 60 && x < NX && x >= 0;
  const int xc = x < 0 ? 0 : (x >= NX ? NX - 1 : x);
  if (k == 0) {
    for (int y = y0; y < y0 + kIpRows && y < NY; y++) {
      if (!xout) continue;
      const long c = (long)x + (long)NX * y;
      idpdx[c] = 0.0;
      idpdy[c] = 0.0;
    }
    return;
  }
  const long o = NXY * (long)k, okm = NXY * (long)d.kmax;
  const bool xint = x >= 2 && x <= NX - 3;
  auto cell = [&](int y) -> long {
    const int yc = y < 0 ? 0 : (y >= NY ? NY - 1 : y);
    return (long)xc + (long)NX * yc;
  };
  // window of rows y-1, y, y+1, y+2 (index 0..3) for b and z; rows y, y+1 of
  // P, sseo and buoy(kmax) are indices 1, 2
  double bw[4], zw[4], pw[4], sw[4], kw[4];
  auto load = [&](int y, int q) {
    const long c = cell(y);
    bw[q] = buoy[o + c];
    zw[q] = zz[o + c];
    pw[q] = P[o + c];
    sw[q] = sseo[c];
    kw[q] = buoy[okm + c];
  };
  load(y0 - 1, 0);
  load(y0, 1);
  load(y0 + 1, 2);
  for (int y = y0; y < y0 + kIpRows; y++) {
    load(y + 2, 3);
    const bool ok = xout && xint && y >= 2 && y <= NY - 3;
    const long c = (long)xc + (long)NX * (y < NY ? y : NY - 1);
    // x: cells x-1..x+2 of row y
    const double b0 = bw[1], z0 = zw[1];
    const double bm = from_lower_lane(b0), zm = from_lower_lane(z0);
    const double b1 = from_upper_lane(b0), z1 = from_upper_lane(z0);
    const double b2 = from_upper_lane(b1), z2 = from_upper_lane(z1);
    const double p1 = from_upper_lane(pw[1]), s1 = from_upper_lane(sw[1]);
    const double k1 = from_upper_lane(kw[1]);
    if (ok && au[c] >= 1) {
      double zrm = z0 - zm, zr0 = z1 - z0, zrp = z2 - z1;
      double brm = b0 - bm, br0 = b1 - b0, brp = b2 - b1;
      double dZx0 = sm_hmx(zrm, zr0), dZx1 = sm_hmx(zr0, zrp);
      double dRx0 = sm_hmx(brm, br0), dRx1 = sm_hmx(br0, brp);
      double AJ = p1 - pw[1];
      double FC = 0.5 * ((b1 + b0) * (z1 - z0) -
                         SM_ONEFIFTH * ((dRx1 - dRx0) * (z1 - z0 - SM_C12 * (dZx1 + dZx0)) -
                                        (dZx1 - dZx0) * (b1 - b0 - SM_C12 * (dRx1 + dRx0))));
      AJ = AJ + FC;
      double h = hun[o + c];
      idpdx[o + c] = AJ * dxm1 * h - h * dxm1 * (s1 - sw[1]) * 0.5 * (k1 + kw[1]);
    }
    if (ok && av[c] >= 1) {
      double zrm = zw[1] - zw[0], zr0 = zw[2] - zw[1], zrp = zw[3] - zw[2];
      double brm = bw[1] - bw[0], br0 = bw[2] - bw[1], brp = bw[3] - bw[2];
      double dZx0 = sm_hmx(zrm, zr0), dZx1 = sm_hmx(zr0, zrp);
      double dRx0 = sm_hmx(brm, br0), dRx1 = sm_hmx(br0, brp);
      double AJ = pw[2] - pw[1];
      double FC = 0.5 * ((bw[2] + bw[1]) * (zw[2] - zw[1]) -
                         SM_ONEFIFTH * ((dRx1 - dRx0) * (zw[2] - zw[1] - SM_C12 * (dZx1 + dZx0)) -
                                        (dZx1 - dZx0) * (bw[2] - bw[1] - SM_C12 * (dRx1 + dRx0))));
      AJ = AJ + FC;
      double h = hvn[o + c];
      idpdy[o + c] = AJ * dym1 * h - h * dym1 * (sw[2] - sw[1]) * 0.5 * (kw[2] + kw[1]);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      bw[q] = bw[q + 1];
      zw[q] = zw[q + 1];
      pw[q] = pw[q + 1];
      sw[q] = sw[q + 1];
      kw[q] = kw[q + 1];
    }
  }
}

// ---------------------------------------------------------------------------
// a13: serial ghost fill, src/futils/halo_zones.F90:141-187 (mirror=.true.):
// the four 1-wide edge assignments applied in order, each reading its whole
// right-hand side first.  aliased == (f1 == f2) decides what the corners see.
__global__ void k_mirror_3d(Dims d, int nk, double* f1, const double* f2, int aliased) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  int k = blockIdx.z;
  if (ii >= d.NX || jj >= d.NY || k >= nk) return;
  const int il = 2, ih = d.NX - 3, jl = 2, jh = d.NY - 3;
  const long NX = d.NX;
  long o = d.NXY * (long)k;
  // value of f2 after the first two (column) statements, as seen by the row
  // statements when f1 and f2 are one array
  auto col_val = [&](int i, int j) -> double {
    if (aliased) {
      if (i == il - 1) return f2[o + il + NX * j];
      if (i == ih + 1) return f2[o + ih + NX * j];
    }
    return f2[o + i + NX * j];
  };
  double v;
  if (jj == jl - 1) v = col_val(ii, jl);
  else if (jj == jh + 1) v = col_val(ii, jh);
  else if (ii == il - 1) v = f2[o + il + NX * jj];
  else if (ii == ih + 1) v = f2[o + ih + NX * jj];
  else return;
  f1[o + ii + NX * jj] = v;
}

// serial update_2d_halo with mirror (halo_zones.F90:102-120): edges then the
// four corners from the diagonal interior cell
__global__ void k_mirror_2d(Dims d, double* f1, const double* f2, int aliased) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const int il = 2, ih = d.NX - 3, jl = 2, jh = d.NY - 3;
  const long NX = d.NX;
  double v;
  bool ci = ii == il - 1 || ii == ih + 1, cj = jj == jl - 1 || jj == jh + 1;
  if (ci && cj) {
    v = f2[(ii == il - 1 ? il : ih) + NX * (jj == jl - 1 ? jl : jh)];
  } else if (cj) {
    int j = jj == jl - 1 ? jl : jh;
    if (aliased && (ii == il - 1 || ii == ih + 1)) v = f2[(ii == il - 1 ? il : ih) + NX * j];
    else v = f2[ii + NX * j];
  } else if (ci) {
    v = f2[(ii == il - 1 ? il : ih) + NX * jj];
  } else {
    return;
  }
  f1[ii + NX * jj] = v;
}

// TWOD exchange between the tiles of one decomposed domain inside one HBM
// (the message pattern of halo_mpi.F90:1095-1154): every ghost cell that lies
// in a neighbour tile's interior takes that value.  One launch covers all
// tiles: blockIdx.z = k, and a flat thread index walks the ghost ring of
// every tile (ring offsets in ring_off[]).
__global__ void k_tile_exchange(int ntiles, int px, TileSet ts, int nk) {
  const long ring_total = ts.ring_off[ntiles];
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int k = blockIdx.y;
  if (q >= ring_total || k >= nk) return;
  // find the tile (few tiles: linear scan)
  int t = 0;
  while (t + 1 < ntiles && ts.ring_off[t + 1] <= q) t++;
  const TileDesc& T = ts.t[t];
  long r = q - ts.ring_off[t];
  long NX = T.ni + 2 * GT_HALO, NY = T.nj + 2 * GT_HALO;
  // ring enumeration: rows jj in [0,2) and [nj+2, nj+4) full width, then
  // columns ii in [0,2) and [ni+2, ni+4) for the interior rows
  long ii, jj;
  long rows = 2 * GT_HALO * NX;
  if (r < rows) {
    long rr = r / NX;
    ii = r % NX;
    jj = rr < GT_HALO ? rr : T.nj + GT_HALO + (rr - GT_HALO);
  } else {
    long rc = r - rows;
    long cc = rc % (2 * GT_HALO);
    jj = GT_HALO + rc / (2 * GT_HALO);
    ii = cc < GT_HALO ? cc : T.ni + GT_HALO + (cc - GT_HALO);
  }
  int ox = ii < GT_HALO ? -1 : (ii >= T.ni + GT_HALO ? 1 : 0);
  int oy = jj < GT_HALO ? -1 : (jj >= T.nj + GT_HALO ? 1 : 0);
  int tx = t % px, ty = t / px;
  int sx = tx + ox, sy = ty + oy;
  int py = ntiles / px;
  if (sx < 0 || sx >= px || sy < 0 || sy >= py) return;
  const TileDesc& S = ts.t[sy * px + sx];
  long SNX = S.ni + 2 * GT_HALO, SNY = S.nj + 2 * GT_HALO;
  long si = ox < 0 ? ii + S.ni : (ox > 0 ? ii - T.ni : ii);
  long sj = oy < 0 ? jj + S.nj : (oy > 0 ? jj - T.nj : jj);
  T.f[ii + NX * (jj + NY * (long)k)] = S.f[si + SNX * (sj + SNY * (long)k)];
}

// ---------------------------------------------------------------------------
// a15: budgets.  Error-free accumulation: each thread keeps a double-double
// (hi, lo) sum of h*f (exact products via fma) and of h; waves and blocks
// combine double-doubles (dd_add, gt_device.h); one partial per block, then
// one block finishes.

template <typename TF>
__global__ void k_budget_partial(Dims d, const int32_t* __restrict__ az,
                                 const double* __restrict__ h, const TF* __restrict__ f,
                                 double* partial) {
  const long nint = (long)(d.NX - 4) * (d.NY - 4) * d.kmax;
  double mh = 0, ml = 0, vh = 0, vl = 0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nint;
       q += (long)gridDim.x * blockDim.x) {
    long ni = d.NX - 4, nj = d.NY - 4;
    long i = q % ni, r = q / ni;
    long j = r % nj, k = 1 + r / nj;
    long c2 = (i + 2) + (long)d.NX * (j + 2);
    if (az[c2] != 1) continue;
    long c = c2 + d.NXY * k;
    double hv = h[c];
    double fv = f ? (double)f[c] : 1.0;
    double p = hv * fv;
    double pe = fma(hv, fv, -p);
    dd_add(mh, ml, p);
    dd_add(mh, ml, pe);
    dd_add(vh, vl, hv);
  }
  __shared__ double sm[4][256];
  int t = threadIdx.x;
  sm[0][t] = mh; sm[1][t] = ml; sm[2][t] = vh; sm[3][t] = vl;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      double a = sm[0][t], b = sm[1][t], c = sm[2][t], e = sm[3][t];
      dd_add_dd(a, b, sm[0][t + s], sm[1][t + s]);
      dd_add_dd(c, e, sm[2][t + s], sm[3][t + s]);
      sm[0][t] = a; sm[1][t] = b; sm[2][t] = c; sm[3][t] = e;
    }
    __syncthreads();
  }
  if (t == 0)
    for (int q = 0; q < 4; q++) partial[4 * blockIdx.x + q] = sm[q][0];
}

__global__ void k_budget_final(int nblocks, const double* __restrict__ partial, double area,
                               double* out) {
  // one block: a strided double-double sum per thread, then a tree
  __shared__ double sm[4][256];
  const int t = threadIdx.x;
  double mh = 0, ml = 0, vh = 0, vl = 0;
  for (int b = t; b < nblocks; b += blockDim.x) {
    dd_add_dd(mh, ml, partial[4 * b], partial[4 * b + 1]);
    dd_add_dd(vh, vl, partial[4 * b + 2], partial[4 * b + 3]);
  }
  sm[0][t] = mh; sm[1][t] = ml; sm[2][t] = vh; sm[3][t] = vl;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      double a = sm[0][t], b = sm[1][t], c = sm[2][t], e = sm[3][t];
      dd_add_dd(a, b, sm[0][t + s], sm[1][t + s]);
      dd_add_dd(c, e, sm[2][t + s], sm[3][t + s]);
      sm[0][t] = a; sm[1][t] = b; sm[2][t] = c; sm[3][t] = e;
    }
    __syncthreads();
  }
  if (t != 0) return;
  mh = sm[0][0]; ml = sm[1][0]; vh = sm[2][0]; vl = sm[3][0];
  double p = mh * area, pe = fma(mh, area, -p);
  pe += ml * area;
  out[0] = p + pe;
  p = vh * area;
  pe = fma(vh, area, -p);
  pe += vl * area;
  out[1] = p + pe;
}

// The budgets of ntr tracers sharing one h in one pass: blockIdx.y is a
// group of up to kBudgetGroup tracers, so h is read once per group.  Each
// thread visits the same cells in the same order as k_budget_partial and the
// block tree is the same, so every tracer's sums (and the volume, formed by
// group 0) are bit for bit those of k_budget_partial.  Partial q (tracer q,
// or the volume at q = ntr) of block b: partial[2*(q*nblocks + b) + {0,1}].
constexpr int kBudgetGroup = 6;

template <typename TF>
__global__ __launch_bounds__(256) void k_budget_batch_partial(
    Dims d, const int32_t* __restrict__ az, int ntr, long tstride,
    const double* __restrict__ h, const TF* __restrict__ f, double* partial) {
  constexpr int G = kBudgetGroup;
  const long nint = (long)(d.NX - 4) * (d.NY - 4) * d.kmax;
  const int t0 = blockIdx.y * G;
  const int ng = min(G, ntr - t0);
  const bool vol = blockIdx.y == 0;
  double mh[G], ml[G], vh = 0, vl = 0;
#pragma unroll
  for (int g = 0; g < G; g++) mh[g] = ml[g] = 0;
  const TF* fg = f + (long)t0 * tstride;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < nint;
       q += (long)gridDim.x * blockDim.x) {
    const long ni = d.NX - 4, nj = d.NY - 4;
    const long i = q % ni, r = q / ni;
    const long j = r % nj, k = 1 + r / nj;
    const long c2 = (i + 2) + (long)d.NX * (j + 2);
    if (az[c2] != 1) continue;
    const long c = c2 + d.NXY * k;
    const double hv = h[c];
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (g < ng) {
        const double fv = (double)fg[(long)g * tstride + c];
        const double p = hv * fv;
        const double pe = fma(hv, fv, -p);
        dd_add(mh[g], ml[g], p);
        dd_add(mh[g], ml[g], pe);
      }
    }
    if (vol) dd_add(vh, vl, hv);
  }
  // one (hi, lo) pair per output and thread, reduced by the tree of
  // k_budget_partial
  __shared__ double sm[2 * (G + 1)][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int g = 0; g < G; g++) {
    sm[2 * g][t] = mh[g];
    sm[2 * g + 1][t] = ml[g];
  }
  sm[2 * G][t] = vh;
  sm[2 * G + 1][t] = vl;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
#pragma unroll
      for (int g = 0; g <= G; g++) {
        double a = sm[2 * g][t], b = sm[2 * g + 1][t];
        dd_add_dd(a, b, sm[2 * g][t + s], sm[2 * g + 1][t + s]);
        sm[2 * g][t] = a;
        sm[2 * g + 1][t] = b;
      }
    }
    __syncthreads();
  }
  const long nb = gridDim.x;
  if (t < ng) {
    double* p = partial + 2 * ((long)(t0 + t) * nb + blockIdx.x);
    p[0] = sm[2 * t][0];
    p[1] = sm[2 * t + 1][0];
  } else if (vol && t == G) {
    double* p = partial + 2 * ((long)ntr * nb + blockIdx.x);
    p[0] = sm[2 * G][0];
    p[1] = sm[2 * G + 1][0];
  }
}

// one block: the npairs (hi, lo) pairs at pq in k_budget_final's order and
// rounding; *outq = sum * area
__device__ __forceinline__ void budget_pairs_block(long npairs, const double* __restrict__ pq,
                                                   double area, double* outq) {
  __shared__ double sm[2][256];
  const int t = threadIdx.x;
  double sh = 0, sl = 0;
  for (long b = t; b < npairs; b += blockDim.x) dd_add_dd(sh, sl, pq[2 * b], pq[2 * b + 1]);
  sm[0][t] = sh;
  sm[1][t] = sl;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      double a = sm[0][t], b = sm[1][t];
      dd_add_dd(a, b, sm[0][t + s], sm[1][t + s]);
      sm[0][t] = a;
      sm[1][t] = b;
    }
    __syncthreads();
  }
  if (t != 0) return;
  const double p = sm[0][0] * area;
  double pe = fma(sm[0][0], area, -p);
  pe += sm[1][0] * area;
  *outq = p + pe;
}

// one block per output q: out[q] from the partials of all blocks
__global__ __launch_bounds__(256) void k_budget_batch_final(int nblocks,
                                                            const double* __restrict__ partial,
                                                            double area, double* out) {
  budget_pairs_block(nblocks, partial + 2 * (long)blockIdx.x * nblocks, area,
                     out + blockIdx.x);
}

// the budgets of several domains in one launch: nout blocks per domain, one
// per output
__global__ __launch_bounds__(256) void k_budget_pairs_final(BudgetPairsSet S) {
  int i = 0;
  long b0 = 0;
  while (i + 1 < S.n && blockIdx.x >= b0 + S.d[i].nout) {
    b0 += S.d[i].nout;
    i++;
  }
  const BudgetPairs& D = S.d[i];
  const long q = blockIdx.x - b0;
  budget_pairs_block(D.npairs, D.partial + 2 * q * D.npairs, D.area, D.out + q);
}

}  // namespace gt

// ---------------------------------------------------------------------------
// host launchers
namespace gt {

static inline dim3 grid2(const Dims& d, int nz, dim3 b) {
  return dim3((d.NX + b.x - 1) / b.x, (d.NY + b.y - 1) / b.y, nz);
}
static const dim3 kBlk2(64, 4, 1);

hipError_t launch_split_u(hipStream_t st, Dims d, int nslice, const double* f,
                          const double* fi_src, double* fi_dst, double* Di, double* adv,
                          const double* U, const double* DU, const uint8_t* mflux,
                          const uint8_t* mupd, double dt, double splitfac, int scheme,
                          double AH, double dx, double dy, double ard1) {
  hipLaunchKernelGGL(k_split_u, grid2(d, nslice, kBlk2), kBlk2, 0, st, d, nslice, f, fi_src,
                     fi_dst, Di, adv, U, DU, mflux, mupd, dt, splitfac, scheme, AH, dx, dy,
                     ard1);
  return hipGetLastError();
}

hipError_t launch_split_v(hipStream_t st, Dims d, int nslice, const double* f,
                          const double* fi_src, double* fi_dst, double* Di, double* adv,
                          const double* V, const double* DV, const uint8_t* mflux,
                          const uint8_t* mupd, double dt, double splitfac, int scheme,
                          double AH, double dx, double dy, double ard1) {
  hipLaunchKernelGGL(k_split_v, grid2(d, nslice, kBlk2), kBlk2, 0, st, d, nslice, f, fi_src,
                     fi_dst, Di, adv, V, DV, mflux, mupd, dt, splitfac, scheme, AH, dx, dy,
                     ard1);
  return hipGetLastError();
}

hipError_t launch_split_w(hipStream_t st, Dims d, const double* f, const double* fi_src,
                          double* fi_dst, double* hi, double* adv, const double* ww,
                          const int32_t* az, double dt, double splitfac, int scheme,
                          int kshift) {
  hipLaunchKernelGGL(k_split_w, grid2(d, 1, kBlk2), kBlk2, 0, st, d, f, fi_src, fi_dst, hi,
                     adv, ww, az, dt, splitfac, scheme, kshift);
  return hipGetLastError();
}

hipError_t launch_split_w_iter(hipStream_t st, Dims d, const double* f, const double* fi_src,
                               double* fi_dst, double* hi, double* adv, const double* ww,
                               const int32_t* az, double* scratch7, unsigned long long* nwarn,
                               double dt, double splitfac, int scheme, int kshift,
                               int itersmax) {
  hipLaunchKernelGGL(k_split_w_iter, grid2(d, 1, kBlk2), kBlk2, 0, st, d, f, fi_src, fi_dst,
                     hi, adv, ww, az, scratch7, nwarn, dt, splitfac, scheme, kshift, itersmax);
  return hipGetLastError();
}

hipError_t launch_finalise(hipStream_t st, Dims d, int k0, int nk, double* f, const double* ho,
                           const double* adv, const double* hi, const uint8_t* mfin,
                           double dt) {
  hipLaunchKernelGGL(k_finalise, grid2(d, nk, kBlk2), kBlk2, 0, st, d, k0, nk, f, ho, adv, hi,
                     mfin, dt);
  return hipGetLastError();
}

hipError_t launch_upstream_2dh(hipStream_t st, Dims d, int nslice, const double* f,
                               const double* fi_src, double* fi_dst, double* Di, double* adv,
                               const double* U, const double* V, const double* Dn,
                               const int32_t* az, double dt, double AH, double dx, double dy,
                               double ard1) {
  hipLaunchKernelGGL(k_upstream_2dh, grid2(d, nslice, kBlk2), kBlk2, 0, st, d, nslice, f,
                     fi_src, fi_dst, Di, adv, U, V, Dn, az, dt, AH, dx, dy, ard1);
  return hipGetLastError();
}

hipError_t launch_eos_linear(hipStream_t st, Dims d, const double* T, const double* S,
                             double* rho, double* buoy, double T0, double S0, double dtr0,
                             double dsr0) {
  long n = d.NXY * (long)d.kmax;
  hipLaunchKernelGGL(k_eos_linear, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     d.NXY, T, S, rho, buoy, T0, S0, dtr0, dsr0);
  return hipGetLastError();
}

hipError_t launch_ip_bm(hipStream_t st, Dims d, const int32_t* az, const int32_t* au,
                        const int32_t* av, const double* H, const double* hn,
                        const double* hun, const double* hvn, const double* buoy, double* zz,
                        double* idpdx, double* idpdy, double dx, double dy) {
  const double dxm1 = 1.0 / dx, dym1 = 1.0 / dy;
  hipLaunchKernelGGL(k_bm_zz, grid2(d, 1, kBlk2), kBlk2, 0, st, d, az, H, hn, zz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bm_idpd, grid2(d, 1, kBlk2), kBlk2, 0, st, d, au, av, buoy, zz, hun, hvn,
                     idpdx, idpdy, dxm1, dym1);
  return hipGetLastError();
}

hipError_t launch_ip_smcw(hipStream_t st, Dims d, const int32_t* az, const int32_t* au,
                          const int32_t* av, const double* H, const double* hn,
                          const double* hun, const double* hvn, const double* buoy,
                          const double* sseo, double* zz, double* P, double* idpdx,
                          double* idpdy, double dx, double dy) {
  const double dxm1 = 1.0 / dx, dym1 = 1.0 / dy;
  hipLaunchKernelGGL(k_sm_column, grid2(d, 1, kBlk2), kBlk2, 0, st, d, az, H, hn, buoy, zz, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int nsx = (d.NX + 59) / 60, nsy = (d.NY + kIpRows - 1) / kIpRows;
  const long nitems = (long)nsx * nsy * (d.kmax + 1);
  hipLaunchKernelGGL(k_sm_idpd_strip, dim3((unsigned)((nitems + 3) / 4)), dim3(256), 0, st, d,
                     nsx, nsy, au, av, buoy, zz, P, hun, hvn, sseo, idpdx, idpdy, dxm1, dym1);
  return hipGetLastError();
}

hipError_t launch_mirror_3d(hipStream_t st, Dims d, double* f1, const double* f2) {
  hipLaunchKernelGGL(k_mirror_3d, grid2(d, d.kmax + 1, kBlk2), kBlk2, 0, st, d, d.kmax + 1, f1,
                     f2, (int)(f1 == f2));
  return hipGetLastError();
}

hipError_t launch_mirror_2d(hipStream_t st, Dims d, double* f1, const double* f2) {
  hipLaunchKernelGGL(k_mirror_2d, grid2(d, 1, kBlk2), kBlk2, 0, st, d, f1, f2,
                     (int)(f1 == f2));
  return hipGetLastError();
}

hipError_t launch_tile_exchange(hipStream_t st, int ntiles, int px, const TileSet& tiles,
                                int nk) {
  const long ring_total = tiles.ring_off[ntiles];
  if (ring_total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tile_exchange, dim3((unsigned)((ring_total + 255) / 256), nk), dim3(256),
                     0, st, ntiles, px, tiles, nk);
  return hipGetLastError();
}

// SUBSTR_INI_PRESS, internal_pressure.F90:235-258: i = imin..imax, j =
// jmin..jmax, k = 0..kmax
__global__ void k_substr_ini_press(Dims d, int first, double* __restrict__ idpdx,
                                   double* __restrict__ idpdy, double* __restrict__ x0,
                                   double* __restrict__ y0) {
  const int i = GT_HALO + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = GT_HALO + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > d.NX - 1 - GT_HALO || j > d.NY - 1 - GT_HALO || k > d.kmax) return;
  const long c = i + (long)d.NX * j + d.NXY * (long)k;
  if (first) {
    x0[c] = idpdx[c];
    idpdx[c] = 0.0;
    y0[c] = idpdy[c];
    idpdy[c] = 0.0;
  } else {
    idpdx[c] = idpdx[c] - x0[c];
    idpdy[c] = idpdy[c] - y0[c];
  }
}

hipError_t launch_substr_ini_press(hipStream_t st, Dims d, int first, double* idpdx,
                                   double* idpdy, double* x0, double* y0) {
  const int ni = d.NX - 2 * GT_HALO, nj = d.NY - 2 * GT_HALO;
  dim3 blk(64, 4);
  dim3 grd((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)(d.kmax + 1));
  hipLaunchKernelGGL(k_substr_ini_press, grd, blk, 0, st, d, first, idpdx, idpdy, x0, y0);
  return hipGetLastError();
}

template <typename TF>
hipError_t launch_budget(hipStream_t st, Dims d, const int32_t* az, const double* h,
                         const TF* f, double* partial, int nblocks, double area, double* out) {
  hipLaunchKernelGGL(k_budget_partial<TF>, dim3(nblocks), dim3(256), 0, st, d, az, h, f, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_budget_final, dim3(1), dim3(256), 0, st, nblocks, partial, area, out);
  return hipGetLastError();
}

template <typename TF>
hipError_t launch_budget_batch(hipStream_t st, Dims d, const int32_t* az, int ntr, long tstride,
                               const double* h, const TF* f, double* partial, int nblocks,
                               double area, double* out) {
  const unsigned ngroups = (unsigned)((ntr + kBudgetGroup - 1) / kBudgetGroup);
  hipLaunchKernelGGL(k_budget_batch_partial<TF>, dim3(nblocks, ngroups), dim3(256), 0, st, d, az,
                     ntr, tstride, h, f, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_budget_batch_final, dim3(ntr + 1), dim3(256), 0, st, nblocks, partial,
                     area, out);
  return hipGetLastError();
}

hipError_t launch_budget_pairs_final(hipStream_t st, const BudgetPairsSet& S) {
  long nblocks = 0;
  for (int i = 0; i < S.n; i++) nblocks += S.d[i].nout;
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_budget_pairs_final, dim3((unsigned)nblocks), dim3(256), 0, st, S);
  return hipGetLastError();
}

}  // namespace gt

namespace gt {
template hipError_t launch_budget<double>(hipStream_t, Dims, const int32_t*, const double*,
                                          const double*, double*, int, double, double*);
template hipError_t launch_budget<float>(hipStream_t, Dims, const int32_t*, const double*,
                                         const float*, double*, int, double, double*);
template hipError_t launch_budget_batch<double>(hipStream_t, Dims, const int32_t*, int, long,
                                                const double*, const double*, double*, int,
                                                double, double*);
template hipError_t launch_budget_batch<float>(hipStream_t, Dims, const int32_t*, int, long,
                                               const double*, const float*, double*, int, double,
                                               double*);
}  // namespace gt
