// gt_eos.h -- the density functions of the equation-of-state module
// (src/3d/eqstate.F90:280-562), shared by the column kernels of do_eqstate
// (gt_eqstate.hip) and the kernels that form the buoyancy inside the
// pressure gradients (gt_ip_fused.hip) so both round identically.  Literals
// without a kind in the reference are FP32 and enter the FP64 arithmetic
// rounded; expression order is the reference's (the library is built with
// -ffp-contract=off).
#pragma once

#include "gt_device.h"

namespace gt {

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

// the buoyancy of do_eqstate, x (rho - rho_0) with x = -g/rho_0
// (eqstate.F90:252-261; g and rho_0 are single-precision literals,
// parameters.F90:21-22)
__device__ __forceinline__ double buoy_of_rho(double r) {
  const double g = (double)9.81f, rho_0 = (double)1025.f;
  const double x = -g / rho_0;
  return x * (r - rho_0);
}

// the density of a water cell by do_eqstate method 2, UNESCO 1980
// (eqstate.F90:151-158), or 3, Jackett et al. at p = 0 with S < 0 taken as 0
// under NONNEGSALT (:197-222)
template <int METHOD>
__device__ __forceinline__ double eqstate_rho(double T, double S, bool nonneg) {
  static_assert(METHOD == 2 || METHOD == 3, "methods 2 and 3 only");
  if constexpr (METHOD == 2) {
    return rho_unesco80(T, S, nonneg);
  } else {
    double s1 = S, r, densp;
    if (nonneg && s1 < 0.0) s1 = 0.0;
    rho_from_theta(s1, T, 0.0, r, densp);
    return r;
  }
}

}  // namespace gt
