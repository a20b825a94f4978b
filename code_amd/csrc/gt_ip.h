// gt_ip.h -- device helpers of the internal-pressure kernels: the linear
// equation of state and the harmonic means of the Shchepetkin-McWilliams
// scheme, shared by the per-method kernels (gt_kernels.hip) and the kernels
// that form the buoyancy inside the pressure gradients (gt_ip_fused.hip) so
// both round identically.
#pragma once

#include "gt_device.h"

namespace gt {

// The reference's eps, OneFifth and 1./12. are single-precision literals
// (src/3d/ip_shchepetkin_mcwilliams.F90:36-37,56-57,139,141)
#define SM_EPS ((double)1.e-10f)
#define SM_ONEFIFTH ((double)0.2f)
#define SM_C12 ((double)(1.0f / 12.0f))

// vertical harmonic mean of the buoyancy differences with the eps test
// (ip_shchepetkin_mcwilliams.F90:108-121)
__device__ __forceinline__ double sm_hmR(double a, double b) {
  double cff = 2.0 * a * b;
  return cff > SM_EPS ? cff / (a + b) : 0.0;
}
// vertical harmonic mean of the height differences (no eps test)
__device__ __forceinline__ double sm_hmZ(double a, double b) { return 2.0 * a * b / (a + b); }
// horizontal harmonic mean with the eps test (:171-188)
__device__ __forceinline__ double sm_hmx(double a, double b) {
  double cff = 2.0 * a * b;
  if (cff > SM_EPS) {
    double cff1 = 1.0 / (a + b);
    return cff * cff1;
  }
  return 0.0;
}

// the linear equation of state and the buoyancy of one cell
// (src/3d/eqstate.F90:127-149,252-261; g and rho_0 are single-precision
// literals, parameters.F90:21-22)
__device__ __forceinline__ double eos_linear_buoy(double T, double S, double T0, double S0,
                                                  double dtr0, double dsr0) {
  const double g = (double)9.81f, rho_0 = (double)1025.f;
  const double x = -g / rho_0;
  const double r = rho_0 + dtr0 * (T - T0) + dsr0 * (S - S0);
  return x * (r - rho_0);
}

}  // namespace gt
