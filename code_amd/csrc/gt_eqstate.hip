// gt_eqstate.hip -- the full equation of state of do_eqstate
// (src/3d/eqstate.F90:94-562): UNESCO 1980 and Jackett et al. (2006)
// density, the interface expansion coefficients alpha/beta, and the buoyancy.
//
// One thread per water column, marching down from the surface: the
// expansion coefficients need the pressure accumulated from the top
// (p = sum of the layer heights above the interface, eqstate.F90:233-240),
// and the interface means of T and S reuse the level above from registers.
// Lanes run along i, so every level is one coalesced row per wavefront.
// Literals without a kind in the reference are FP32 and enter the FP64
// arithmetic rounded; expression order is the reference's (the library is
// built with -ffp-contract=off), so results equal the CPU build's bit for bit.
#include "gt_eos.h"
#include "gt_kernels.h"

namespace gt {

namespace {

constexpr int kEosBlock = 256;
constexpr int kEosUnroll = 4;

// do_eqstate methods 2 and 3 (eqstate.F90:151-245) and the buoyancy
// (:252-261) for one column per thread.  Land columns keep rho and only get
// buoy.  AB: also the expansion coefficients (beta = d rho/dS, alpha =
// d rho/d theta at the interfaces k = kmax-1..1; zero at kmax).
template <int METHOD, bool NONNEG, bool AB>
__global__ void __launch_bounds__(kEosBlock)
    k_eqstate_column(long NXY, int kmax, const int32_t* __restrict__ az,
                     const double* __restrict__ T, const double* __restrict__ S,
                     const double* __restrict__ hn, double* __restrict__ rho,
                     double* __restrict__ buoy, double* __restrict__ alpha,
                     double* __restrict__ beta) {
  const long c = blockIdx.x * (long)kEosBlock + threadIdx.x;
  if (c >= NXY) return;
  const bool wet = az[c] > 0;
  double Tup = 0.0, Sup = 0.0, hup = 0.0, p1 = 0.0;
  int k = kmax;
  while (k >= 1) {
    const int nk = k >= kEosUnroll ? kEosUnroll : k;
    double t[kEosUnroll], s[kEosUnroll], h[kEosUnroll], r[kEosUnroll];
    // issue every load of the chunk before any arithmetic
#pragma unroll
    for (int u = 0; u < kEosUnroll; u++) {
      if (u < nk) {
        const long q = c + NXY * (long)(k - u);
        if (wet) {
          t[u] = T[q];
          s[u] = S[q];
          if (AB) h[u] = hn[q];
        } else {
          r[u] = rho[q];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kEosUnroll; u++) {
      if (u < nk) {
        const int kk = k - u;
        const long q = c + NXY * (long)kk;
        if (wet) {
          r[u] = eqstate_rho<METHOD>(t[u], s[u], NONNEG);
          rho[q] = r[u];
          if (AB) {
            if (kk == kmax) {
              alpha[q] = 0.0;
              beta[q] = 0.0;
            } else {
              const double th = 0.5 * (Tup + t[u]);
              const double s1 = 0.5 * (Sup + s[u]);
              p1 = p1 + hup;
              double rs, rt;
              eosall_from_theta(s1, th, p1, rs, rt);
              beta[q] = rs;
              alpha[q] = rt;
            }
            Tup = t[u];
            Sup = s[u];
            hup = h[u];
          }
        }
        buoy[q] = buoy_of_rho(r[u]);
      }
    }
    k -= nk;
  }
}

// do_eqstate method 1, the whole-array assignments alpha = dtr0, beta = dsr0
// (eqstate.F90:139-147)
__global__ void k_fill2(long n, double* __restrict__ a, double va, double* __restrict__ b,
                        double vb) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (a) a[q] = va;
  if (b) b[q] = vb;
}

// the three point functions of the module, elementwise
__global__ void k_eos_points(int fn, long n, const double* __restrict__ s,
                             const double* __restrict__ th, const double* __restrict__ p,
                             double* __restrict__ o1, double* __restrict__ o2, int nonneg) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= n) return;
  if (fn == 0) {
    o1[q] = rho_unesco80(th[q], s[q], nonneg != 0);
  } else if (fn == 1) {
    double a, b;
    rho_from_theta(s[q], th[q], p ? p[q] : 0.0, a, b);
    o1[q] = a;
    if (o2) o2[q] = b;
  } else {
    double a, b;
    eosall_from_theta(s[q], th[q], p ? p[q] : 0.0, a, b);
    o1[q] = a;
    o2[q] = b;
  }
}

// the same without the expansion coefficients: no level depends on another,
// so one thread per cell of levels 1..kmax (a column per thread would leave
// the chip with a few wavefronts per SIMD, each a serial chain of divisions)
template <int METHOD, bool NONNEG>
__global__ void __launch_bounds__(kEosBlock)
    k_eqstate_cells(long NXY, long ncell, const int32_t* __restrict__ az,
                    const double* __restrict__ T, const double* __restrict__ S,
                    double* __restrict__ rho, double* __restrict__ buoy) {
  const long i = blockIdx.x * (long)kEosBlock + threadIdx.x;
  if (i >= ncell) return;
  const long q = NXY + i;  // level 0 is not touched
  const long c = i % NXY;
  double r;
  if (az[c] > 0) {
    r = eqstate_rho<METHOD>(T[q], S[q], NONNEG);
    rho[q] = r;
  } else {
    r = rho[q];
  }
  buoy[q] = buoy_of_rho(r);
}

template <int METHOD, bool NONNEG>
hipError_t launch_column(hipStream_t st, Dims d, const int32_t* az, const double* T,
                         const double* S, const double* hn, double* rho, double* buoy,
                         double* alpha, double* beta) {
  const dim3 grid((unsigned)((d.NXY + kEosBlock - 1) / kEosBlock));
  if (alpha && beta)
    hipLaunchKernelGGL((k_eqstate_column<METHOD, NONNEG, true>), grid, dim3(kEosBlock), 0, st,
                       d.NXY, d.kmax, az, T, S, hn, rho, buoy, alpha, beta);
  else {
    const long ncell = d.NXY * (long)d.kmax;
    hipLaunchKernelGGL((k_eqstate_cells<METHOD, NONNEG>),
                       dim3((unsigned)((ncell + kEosBlock - 1) / kEosBlock)), dim3(kEosBlock), 0,
                       st, d.NXY, ncell, az, T, S, rho, buoy);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_eqstate(hipStream_t st, Dims d, int method, bool nonnegsalt,
                          const int32_t* az, const double* T, const double* S,
                          const double* hn, double* rho, double* buoy, double* alpha,
                          double* beta, double T0, double S0, double dtr0, double dsr0) {
  if (method == 1) {
    hipError_t e = launch_eos_linear(st, d, T, S, rho, buoy, T0, S0, dtr0, dsr0);
    if (e != hipSuccess || (!alpha && !beta)) return e;
    const long n = d.NXY * (long)(d.kmax + 1);
    hipLaunchKernelGGL(k_fill2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, alpha,
                       dtr0, beta, dsr0);
    return hipGetLastError();
  }
  if (method == 2)
    return nonnegsalt ? launch_column<2, true>(st, d, az, T, S, hn, rho, buoy, alpha, beta)
                      : launch_column<2, false>(st, d, az, T, S, hn, rho, buoy, alpha, beta);
  return nonnegsalt ? launch_column<3, true>(st, d, az, T, S, hn, rho, buoy, alpha, beta)
                    : launch_column<3, false>(st, d, az, T, S, hn, rho, buoy, alpha, beta);
}

hipError_t launch_eos_points(hipStream_t st, int fn, long n, const double* s, const double* th,
                             const double* p, double* o1, double* o2, bool nonnegsalt) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_eos_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, fn, n,
                     s, th, p, o1, o2, nonnegsalt ? 1 : 0);
  return hipGetLastError();
}

}  // namespace gt
