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
#include "gt_kernels.h"

namespace gt {

namespace {

// rho_from_theta_unesco80, eqstate.F90:280-329
__device__ __forceinline__ double rho_unesco80(double T, double S, bool nonneg) {
  const double t1 = T, t2 = t1 * t1, t3 = t1 * t2, t4 = t2 * t2, t5 = t1 * t4;
  double s1 = S;
  if (nonneg && s1 < 0.0) s1 = 0.0;
  const double s2 = s1 * s1, s3 = s1 * s2;
  double x = (double)999.842594f + (double)6.793952e-02f * t1 - (double)9.09529e-03f * t2 +
             (double)1.001685e-04f * t3;
  x = x - (double)1.120083e-06f * t4 + (double)6.536332e-09f * t5;
  x = x + s1 * ((double)0.824493f - (double)4.0899e-03f * t1 + (double)7.6438e-05f * t2 -
                (double)8.2467e-07f * t3);
  x = x + s1 * (double)5.3875e-09f * t4;
  x = x + sqrt(s3) * ((double)-5.72466e-03f + (double)1.0227e-04f * t1 -
                      (double)1.6546e-06f * t2);
  x = x + (double)4.8314e-04f * s2;
  return x;
}

// the Jackett et al. (2006) rational function at p = 0, eqstate.F90:381-400
struct Jackett {
  double anum, aden;
};
__device__ __forceinline__ Jackett jackett_p0(double s, double th, double th2, double sqrts) {
  Jackett j;
  j.anum = 9.9984085444849347e+02 +
           th * (7.3471625860981584e+00 +
                 th * (-5.3211231792841769e-02 + th * 3.6492439109814549e-04)) +
           s * (2.5880571023991390e+00 - th * 6.7168282786692355e-03 +
                s * 1.9203202055760151e-03);
  j.aden = 1.0000000000000000e+00 +
           th * (7.2815210113327091e-03 +
                 th * (-4.4787265461983921e-05 +
                       th * (3.3851002965802430e-07 + th * 1.3651202389758572e-10))) +
           s * (1.7632126669040377e-03 -
                th * (8.8066583251206474e-06 + th2 * 1.8832689434804897e-10) +
                sqrts * (5.7463776745432097e-06 + th2 * 1.4716275472242334e-09));
  return j;
}

// pressure terms, eqstate.F90:407-417
__device__ __forceinline__ void jackett_p(Jackett& j, double s, double th2, double p,
                                          double pth) {
  j.anum = j.anum + p * (1.1798263740430364e-02 + th2 * 9.8920219266399117e-08 +
                         s * 4.6996642771754730e-06 -
                         p * (2.5862187075154352e-08 + th2 * 3.2921414007960662e-12));
  j.aden = j.aden + p * (6.7103246285651894e-06 -
                         pth * (th2 * 2.4461698007024582e-17 + p * 9.1534417604289062e-18));
}

// rho_from_theta, eqstate.F90:338-423
__device__ __forceinline__ void rho_from_theta(double s, double th, double p, double& dens0,
                                               double& densp) {
  const double th2 = th * th, sqrts = sqrt(s);
  Jackett j = jackett_p0(s, th, th2, sqrts);
  dens0 = j.anum / j.aden;
  densp = dens0;
  if (p != 0.0) {
    jackett_p(j, s, th2, p, p * th);
    densp = j.anum / j.aden;
  }
}

// eosall_from_theta, eqstate.F90:432-562
__device__ __forceinline__ void eosall_from_theta(double s, double th, double p, double& rho_s,
                                                  double& rho_th) {
  const double th2 = th * th, sqrts = sqrt(s);
  Jackett j = jackett_p0(s, th, th2, sqrts);
  double anum_s =
      2.5880571023991390e+00 - th * 6.7168282786692355e-03 + s * 3.8406404111520300e-03;
  const double aden_s = 1.7632126669040377e-03 +
                        th * (-8.8066583251206470e-06 - th2 * 1.8832689434804897e-10) +
                        sqrts * (8.6195665118148150e-06 + th2 * 2.2074413208363504e-09);
  double anum_th = 7.3471625860981580e+00 +
                   th * (-1.0642246358568354e-01 + th * 1.0947731732944364e-03) -
                   s * 6.7168282786692355e-03;
  double aden_th = 7.2815210113327090e-03 +
                   th * (-8.9574530923967840e-05 +
                         th * (1.0155300889740728e-06 + th * 5.4604809559034290e-10)) +
                   s * (-8.8066583251206470e-06 - th2 * 5.6498068304414700e-10 +
                        th * sqrts * 2.9432550944484670e-09);
  if (p != 0.0) {
    const double pth = p * th;
    jackett_p(j, s, th2, p, pth);
    anum_s = anum_s + p * 4.6996642771754730e-06;
    anum_th = anum_th + pth * (1.9784043853279823e-07 - p * 6.5842828015921320e-12);
    aden_th = aden_th - p * p * (th2 * 7.3385094021073750e-17 + p * 9.1534417604289060e-18);
  }
  const double rec_aden = 1.0 / j.aden;
  const double rho = j.anum * rec_aden;
  rho_s = (anum_s - aden_s * rho) * rec_aden;
  rho_th = (anum_th - aden_th * rho) * rec_aden;
}

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
  const double g = (double)9.81f, rho_0 = (double)1025.f;
  const double x = -g / rho_0;
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
          if (METHOD == 2) {
            r[u] = rho_unesco80(t[u], s[u], NONNEG);
          } else {
            double s1 = s[u], densp;
            if (NONNEG && s1 < 0.0) s1 = 0.0;
            rho_from_theta(s1, t[u], 0.0, r[u], densp);
          }
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
        buoy[q] = x * (r[u] - rho_0);
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

template <int METHOD, bool NONNEG>
hipError_t launch_column(hipStream_t st, Dims d, const int32_t* az, const double* T,
                         const double* S, const double* hn, double* rho, double* buoy,
                         double* alpha, double* beta) {
  const dim3 grid((unsigned)((d.NXY + kEosBlock - 1) / kEosBlock));
  if (alpha && beta)
    hipLaunchKernelGGL((k_eqstate_column<METHOD, NONNEG, true>), grid, dim3(kEosBlock), 0, st,
                       d.NXY, d.kmax, az, T, S, hn, rho, buoy, alpha, beta);
  else
    hipLaunchKernelGGL((k_eqstate_column<METHOD, NONNEG, false>), grid, dim3(kEosBlock), 0,
                       st, d.NXY, d.kmax, az, T, S, hn, rho, buoy, alpha, beta);
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
