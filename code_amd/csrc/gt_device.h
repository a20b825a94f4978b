// gt_device.h -- device-side helpers shared by the transport kernels.
//
// Field layout is the reference's: f(imin-2:imax+2, jmin-2:jmax+2, 0:kmax),
// i fastest (include/cppdefs.h:146-159 of getm-model/code).  Kernels index
// with ii = i-(imin-2) in [0,NX) and jj = j-(jmin-2) in [0,NY); the interior
// is ii in [2, NX-3].  Every arithmetic expression keeps the reference's
// left-to-right order; the library is built with -ffp-contract=off so the
// FP64 path rounds exactly as the CPU build does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GT_HALO 2

namespace gt {

enum Scheme { NOADV = 0, UPSTREAM = 1, UPSTREAM_2DH = 2, P2 = 3, SUPERBEE = 4,
              MUSCL = 5, P2_PDM = 6, J7 = 7, FCT = 8, P2_2DH = 9 };
enum Split { NOSPLIT = 0, FULLSPLIT = 1, HALFSPLIT = 2, HVSPLIT = 3 };
enum Tag { H_TAG = 10, D_TAG = 20, U_TAG = 31, V_TAG = 32, W_TAG = 33 };

struct Dims {
  int NX, NY, kmax;
  long NXY;
};

// ---- cross-lane moves (DPP wave shifts; call with every lane active) -------
__device__ __forceinline__ int from_lower_lane(int v) {  // lane l gets lane l-1 (0 in lane 0)
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int from_upper_lane(int v) {  // lane l gets lane l+1 (0 in lane 63)
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ float from_lower_lane(float v) {
  return __builtin_bit_cast(float, from_lower_lane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float from_upper_lane(float v) {
  return __builtin_bit_cast(float, from_upper_lane(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double from_lower_lane(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)from_lower_lane((int)(unsigned)b);
  const unsigned hi = (unsigned)from_lower_lane((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}
__device__ __forceinline__ double from_upper_lane(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)from_upper_lane((int)(unsigned)b);
  const unsigned hi = (unsigned)from_upper_lane((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, (unsigned long long)lo | ((unsigned long long)hi << 32));
}

// double-double accumulation (Knuth's two-sum of the running sum and the
// addend, the error carried in sl): the budgets' compensated sums
__device__ __forceinline__ void dd_add(double& sh, double& sl, double a) {
  double s = sh + a;
  double bv = s - sh;
  double e = (sh - (s - bv)) + (a - bv);
  e += sl;
  double r = s + e;
  sl = e - (r - s);
  sh = r;
}
__device__ __forceinline__ void dd_add_dd(double& sh, double& sl, double ah, double al) {
  dd_add(sh, sl, ah);
  dd_add(sh, sl, al);
}

template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }

// adv_interfacial_reconstruction_p2, src/2d/advection.F90:854-879
template <typename T>
__device__ __forceinline__ T reconstruct_p2(T cfl, T fu, T deltafu, T deltaf) {
  const T one6th = T(1) / T(6);
  T x = one6th * (T(1) - T(2) * cfl);
  return fu + T(0.5) * (T(1) - cfl) * ((T(0.5) + x) * deltaf + (T(0.5) - x) * deltafu);
}

// adv_interfacial_reconstruction, src/2d/advection.F90:785-845
template <typename T>
__device__ __forceinline__ T reconstruct(int scheme, T cfl, T fuu, T fu, T fd) {
  const T one6th = T(1) / T(6);
  T deltaf = fd - fu;
  T deltafu = fu - fuu;
  if (deltaf * deltafu > T(0)) {
    T ratio = deltafu / deltaf;
    T limiter = T(0);
    switch (scheme) {
      case P2_PDM: {
        T x = one6th * (T(1) - T(2) * cfl);
        limiter = (T(0.5) + x) + (T(0.5) - x) * ratio;
        limiter = tmin(tmin(T(2) * ratio / (cfl + T(1.e-10)), limiter), T(2) / (T(1) - cfl));
        break;
      }
      case SUPERBEE:
        limiter = tmax(tmin(T(2) * ratio, T(1)), tmin(ratio, T(2)));
        break;
      case MUSCL:
        limiter = tmin(tmin(T(2) * ratio, T(0.5) * (T(1) + ratio)), T(2));
        break;
      case P2:
        return reconstruct_p2(cfl, fu, deltafu, deltaf);
      default:  // UPSTREAM
        return fu;
    }
    return fu + T(0.5) * limiter * (T(1) - cfl) * deltaf;
  }
  if (scheme == P2) return reconstruct_p2(cfl, fu, deltafu, deltaf);
  return fu;
}

}  // namespace gt
