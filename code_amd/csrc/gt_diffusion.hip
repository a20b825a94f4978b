// gt_diffusion.hip -- the step after advection: the implicit vertical
// diffusion of do_temperature and do_salinity (temperature.F90:463-551,
// salinity.F90:346-393) and the batched getm_tridiagonal
// (src/futils/tridiagonal.F90).
//
// One thread per water column.  The elimination runs down from the surface
// (the reference's lt = kmax to fi = 1), forming each row of the
// Crank-Nicolson matrix on the fly from a three-level window of the tracer
// and the layer heights, and keeps the eliminated coefficients ru, qu in
// registers for columns of up to 48 levels (in a level-major work field
// beyond that); the substitution runs back up and writes the tracer.
// Expression order is the reference's (built with -ffp-contract=off):
// bit-identical to the CPU build, except the short-wave exponentials, which
// go through the device exp (within an ulp of libm's).
#include "gt_kernels.h"

namespace gt {

namespace {

constexpr int kBx = 64, kBy = 4;
constexpr int kChunk = 4;  // levels whose loads are issued ahead of their arithmetic

struct DiffArgs {
  Dims d;
  double dt, cnpar, avm;
  const int32_t* az;
  const double* hn;
  const double* nuh;
  double* f;
  double* ru;  // work: level-major, one value per cell
  double* qu;
  // temperature only
  const double *swr, *shf, *A, *g1, *g2, *dry_z, *S;
  double* rad;
  double* rad1d;  // work: the reflected profile (when reflection is on)
  double refl_frac, min_bot_frac, rho_0_cpi;
};

// KM = 0: ru, qu in the level-major work field (any kmax); KM > 0: in
// registers (kmax <= KM), every level's loads independent of the arithmetic
template <bool TEMP, int KM>
__global__ void __launch_bounds__(kBx* kBy) k_vertical_diffusion(DiffArgs a) {
  const int ii = blockIdx.x * kBx + threadIdx.x + 2;
  const int jj = blockIdx.y * kBy + threadIdx.y + 2;
  const Dims d = a.d;
  if (ii > d.NX - 3 || jj > d.NY - 3) return;
  const long c = ii + (long)d.NX * jj;
  const int azc = a.az[c];
  const long NXY = d.NXY;
  const int kmax = d.kmax;
#define L(p, k) (p)[c + NXY * (long)(k)]
  // where the short-wave profile is read from (rad, or the reflected copy)
  const double* radsrc = a.rad;
  double shf_dz = 0.0, dzdt = 0.0;
  if (TEMP) {
    if (azc < 1) return;
    // solar radiation, temperature.F90:466-480
    const double swr = a.swr[c], Ac = a.A[c], g1 = a.g1[c], g2 = a.g2[c];
    L(a.rad, kmax) = swr;
    double zz = 0.0;
    double rad0 = swr;
    // kChunk levels of hn per memory round trip; the levels below 0 of the
    // last chunk are selected away (a select, not a branch, so the loads stay
    // ahead of their uses)
    for (int kb = kmax - 1; kb >= 0; kb -= kChunk) {
      double q[kChunk];
#pragma unroll
      for (int v = 0; v < kChunk; v++) q[v] = L(a.hn, max(kb - v, 0) + 1);
#pragma unroll
      for (int v = 0; v < kChunk; v++) {
        const bool on = kb - v >= 0;
        const double zn = zz + q[v];
        zz = on ? zn : zz;
        const double r = swr * (Ac * exp(-zz / g1) + (1 - Ac) * exp(-zz / g2));
        rad0 = on ? r : rad0;
        if (on) L(a.rad, kb - v) = rad0;
      }
    }
    if (azc != 1) return;
    // the primitive sea-ice model, :490-493
    double shf_loc = a.shf[c];
    if (L(a.f, kmax) <= (double)-0.0575f * L(a.S, kmax)) shf_loc = shf_loc < 0.0 ? 0.0 : shf_loc;
    // bottom reflection, :486-500
    if (a.refl_frac > 0.0 && rad0 / swr > a.min_bot_frac) {
      const double swr_refl = rad0 * a.refl_frac;
      L(a.rad1d, 0) = rad0 - swr_refl;
      double z2 = 0.0;
      for (int kb = 1; kb <= kmax; kb += kChunk) {  // loads ahead, as above
        double qh[kChunk], qr[kChunk];
#pragma unroll
        for (int v = 0; v < kChunk; v++) {
          qh[v] = L(a.hn, min(kb + v, kmax));
          qr[v] = L(a.rad, min(kb + v, kmax));
        }
#pragma unroll
        for (int v = 0; v < kChunk; v++) {
          const bool on = kb + v <= kmax;
          const double zn = z2 + qh[v];
          z2 = on ? zn : z2;
          const double r = qr[v] - swr_refl * (Ac * exp(-z2 / g1) + (1 - Ac) * exp(-z2 / g2));
          if (on) L(a.rad1d, kb + v) = r;
        }
      }
      radsrc = a.rad1d;
    }
    shf_dz = shf_loc * a.rho_0_cpi;
    dzdt = a.dry_z[c] * a.dt;
    if (kmax <= 1) return;
  } else {
    if (azc != 1 || kmax <= 1) return;
  }
  const double dt = a.dt, cnpar = a.cnpar, avm = a.avm, rc = a.rho_0_cpi;
  if constexpr (KM > 0) {
    // rows k = kmax (surface) .. 1 (bottom); carried: f(k+1), f(k), hn(k),
    // auxo/auxn of interface k, the short-wave at k
    double ruv[KM + 1], quv[KM + 1];
    double fkp = 0.0, fk = 0.0, hk = 0.0, auxo_k = 0.0, auxn_k = 0.0, rk = 0.0;
    double ru = 0.0, qu = 0.0;
    const double h_top = L(a.hn, kmax), f_top = L(a.f, kmax);
    const double r_top = TEMP ? L(radsrc, kmax) : 0.0;
    // level k-1 of rows k = kc..kc-kChunk+1, loaded together (one memory
    // round trip per kChunk rows); rows past kmax read level kmax, unused
    double ph[kChunk], pf[kChunk], pn[kChunk], pr[kChunk];
#pragma unroll
    for (int k = KM; k >= 1; k--) {
      const int u = (KM - k) % kChunk;
      if (u == 0) {
#pragma unroll
        for (int v = 0; v < kChunk; v++) {
          const int l = min(max(k - v - 1, 0), kmax);
          ph[v] = L(a.hn, l);
          pf[v] = L(a.f, l);
          pn[v] = L(a.nuh, l);
          if (TEMP) pr[v] = L(radsrc, l);
        }
      }
      // a guard, not continue/break: the loop must unroll fully so ruv/quv
      // stay in registers
      if (k <= kmax) {
      if (k == kmax) {
        hk = h_top;
        fk = f_top;
        if (TEMP) rk = r_top * rc;
      }
      double hkm = 0.0, fkm = 0.0, auxo_m = 0.0, auxn_m = 0.0, rkm = 0.0;
      if (k >= 2) {
        hkm = ph[u];
        fkm = pf[u];
        const double nk = pn[u];
        auxo_m = 2.0 * (1 - cnpar) * dt * (nk + avm) / (hk + hkm);
        auxn_m = 2.0 * cnpar * dt * (nk + avm) / (hk + hkm);
      }
      if (TEMP) rkm = pr[u] * rc;
      if (k == kmax) {
        const double a1 = -auxn_m, a2 = hk + auxn_m;
        double a4 = fk * (hk - auxo_m) + fkm * auxo_m;
        if (TEMP) a4 = a4 + dzdt * (rk + shf_dz - rkm);
        ru = a1 / a2;
        qu = a4 / a2;
      } else if (k >= 2) {
        const double a3 = -auxn_k, a1 = -auxn_m;
        const double a2 = hk + auxn_k + auxn_m;
        double a4 = fkp * auxo_k + fk * (hk - auxo_k - auxo_m) + fkm * auxo_m;
        if (TEMP) a4 = a4 + dzdt * (rk - rkm);
        const double ru_n = a1 / (a2 - a3 * ru);
        qu = (a4 - a3 * qu) / (a2 - a3 * ru);
        ru = ru_n;
      } else {
        const double a3 = -auxn_k, a2 = hk + auxn_k;
        double a4 = fkp * auxo_k + fk * (hk - auxo_k);
        if (TEMP) a4 = a4 + dzdt * (rk - rkm);
        qu = (a4 - a3 * qu) / (a2 - a3 * ru);
      }
      ruv[k] = ru;
      quv[k] = qu;
      fkp = fk;
      fk = fkm;
      hk = hkm;
      auxo_k = auxo_m;
      auxn_k = auxn_m;
      rk = rkm;
      }
    }
    double v = quv[1];
    L(a.f, 1) = v;
#pragma unroll
    for (int k = 2; k <= KM; k++) {
      if (k <= kmax) {
        v = quv[k] - ruv[k] * v;
        L(a.f, k) = v;
      }
    }
    return;
  }
  // the surface row, k = kmax (auxo/auxn of interface kmax-1)
  double hk = L(a.hn, kmax), hkm = L(a.hn, kmax - 1);
  double fk = L(a.f, kmax), fkm = L(a.f, kmax - 1);
  double nk = L(a.nuh, kmax - 1);
  double auxo_m = 2.0 * (1 - cnpar) * dt * (nk + avm) / (hk + hkm);
  double auxn_m = 2.0 * cnpar * dt * (nk + avm) / (hk + hkm);
  double rk = 0.0, rkm = 0.0;
  if (TEMP) {
    rk = L(radsrc, kmax) * rc;
    rkm = L(radsrc, kmax - 1) * rc;
  }
  double a4 = fk * (hk - auxo_m) + fkm * auxo_m;
  if (TEMP) a4 = a4 + dzdt * (rk + shf_dz - rkm);
  double a1 = -auxn_m, a2 = hk + auxn_m;
  double ru = a1 / a2, qu = a4 / a2;
  L(a.ru, kmax) = ru;
  L(a.qu, kmax) = qu;
  // down the column: row k uses interfaces k (carried) and k-1.  The
  // level-(k-1) inputs of four rows are loaded before their arithmetic.
  for (int k0 = kmax - 1; k0 >= 1; k0 -= kChunk) {
    double hv[kChunk], fv[kChunk], nv[kChunk], rv[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; u++) {
      const int k = k0 - u;
      if (k >= 2) {
        hv[u] = L(a.hn, k - 1);
        fv[u] = L(a.f, k - 1);
        nv[u] = L(a.nuh, k - 1);
      }
      if (TEMP && k >= 1) rv[u] = L(radsrc, k - 1);
    }
#pragma unroll
    for (int u = 0; u < kChunk; u++) {
      const int k = k0 - u;
      if (k < 1) break;
      const double auxo_k = auxo_m, auxn_k = auxn_m;
      const double fkp = fk;
      fk = fkm;
      hk = hkm;
      if (TEMP) rk = rkm;
      if (k >= 2) {
        hkm = hv[u];
        fkm = fv[u];
        nk = nv[u];
        auxo_m = 2.0 * (1 - cnpar) * dt * (nk + avm) / (hk + hkm);
        auxn_m = 2.0 * cnpar * dt * (nk + avm) / (hk + hkm);
        if (TEMP) rkm = rv[u] * rc;
        const double a3 = -auxn_k;
        a1 = -auxn_m;
        a2 = hk + auxn_k + auxn_m;
        a4 = fkp * auxo_k + fk * (hk - auxo_k - auxo_m) + fkm * auxo_m;
        if (TEMP) a4 = a4 + dzdt * (rk - rkm);
        const double ru_n = a1 / (a2 - a3 * ru);
        qu = (a4 - a3 * qu) / (a2 - a3 * ru);
        ru = ru_n;
        L(a.ru, k) = ru;
        L(a.qu, k) = qu;
      } else {
        // the bottom row, k = fi = 1
        if (TEMP) rkm = rv[u] * rc;
        const double a3 = -auxn_k;
        a2 = hk + auxn_k;
        a4 = fkp * auxo_k + fk * (hk - auxo_k);
        if (TEMP) a4 = a4 + dzdt * (rk - rkm);
        qu = (a4 - a3 * qu) / (a2 - a3 * ru);
      }
    }
  }
  // back up: value(1) = qu(1), value(k) = qu(k) - ru(k) value(k-1)
  double v = qu;
  L(a.f, 1) = v;
  for (int k0 = 2; k0 <= kmax; k0 += kChunk) {
    double qv[kChunk], rw[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; u++)
      if (k0 + u <= kmax) {
        qv[u] = L(a.qu, k0 + u);
        rw[u] = L(a.ru, k0 + u);
      }
#pragma unroll
    for (int u = 0; u < kChunk; u++)
      if (k0 + u <= kmax) {
        v = qv[u] - rw[u] * v;
        L(a.f, k0 + u) = v;
      }
  }
#undef L
}

// getm_tridiagonal over n systems, level-major: element k of system s at
// [k*n + s]
__global__ void k_tridiagonal(long n, int fi, int lt, const double* __restrict__ au,
                              const double* __restrict__ bu, const double* __restrict__ cu,
                              const double* __restrict__ du, double* __restrict__ value,
                              double* __restrict__ ru, double* __restrict__ qu) {
  const long s = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (s >= n) return;
#define E(p, k) (p)[(long)(k) * n + s]
  double r = E(au, lt) / E(bu, lt);
  double q = E(du, lt) / E(bu, lt);
  E(ru, lt) = r;
  E(qu, lt) = q;
  for (int i = lt - 1; i >= fi + 1; i--) {
    const double c = E(cu, i), b = E(bu, i);
    const double rn = E(au, i) / (b - c * r);
    q = (E(du, i) - c * q) / (b - c * r);
    r = rn;
    E(ru, i) = r;
    E(qu, i) = q;
  }
  q = (E(du, fi) - E(cu, fi) * q) / (E(bu, fi) - E(cu, fi) * r);
  double v = q;
  E(value, fi) = v;
  for (int i = fi + 1; i <= lt; i++) {
    v = E(qu, i) - E(ru, i) * v;
    E(value, i) = v;
  }
#undef E
}

}  // namespace

hipError_t launch_vertical_diffusion(hipStream_t st, Dims d, bool temperature, double dt,
                                     double cnpar, const int32_t* az, const double* hn,
                                     const double* nuh, double* f, double* work,
                                     const double* swr, const double* shf, const double* A,
                                     const double* g1, const double* g2, const double* dry_z,
                                     const double* S, double refl_frac, double min_bot_frac,
                                     double* rad) {
  if (d.NX < 5 || d.NY < 5) return hipSuccess;
  DiffArgs a{};
  a.d = d;
  a.dt = dt;
  a.cnpar = cnpar;
  a.az = az;
  a.hn = hn;
  a.nuh = nuh;
  a.f = f;
  const long N3 = d.NXY * (long)(d.kmax + 1);
  a.ru = work;
  a.qu = work + N3;
  a.rad1d = work + 2 * N3;
  a.swr = swr;
  a.shf = shf;
  a.A = A;
  a.g1 = g1;
  a.g2 = g2;
  a.dry_z = dry_z;
  a.S = S;
  a.rad = rad;
  a.refl_frac = refl_frac;
  a.min_bot_frac = min_bot_frac;
  // parameters.F90:22-27 (FP32 literals)
  a.rho_0_cpi = 1.0 / ((double)1025.f * (double)3985.f);
  a.avm = temperature ? (double)1.4e-7f : (double)1.1e-9f;
  const dim3 g((unsigned)((d.NX - 4 + kBx - 1) / kBx), (unsigned)((d.NY - 4 + kBy - 1) / kBy));
  // the smallest register variant that holds the column, else the work field
#define GT_VD(KM)                                                                  \
  do {                                                                             \
    if (temperature)                                                               \
      hipLaunchKernelGGL((k_vertical_diffusion<true, KM>), g, dim3(kBx, kBy), 0, st, a); \
    else                                                                           \
      hipLaunchKernelGGL((k_vertical_diffusion<false, KM>), g, dim3(kBx, kBy), 0, st, a); \
  } while (0)
  if (d.kmax <= 16)
    GT_VD(16);
  else if (d.kmax <= 32)
    GT_VD(32);
  else if (d.kmax <= 48)
    GT_VD(48);
  else
    GT_VD(0);
#undef GT_VD
  return hipGetLastError();
}

hipError_t launch_tridiagonal(hipStream_t st, long n, int fi, int lt, const double* au,
                              const double* bu, const double* cu, const double* du,
                              double* value, double* work) {
  if (n <= 0) return hipSuccess;
  const long nk = n * (long)(lt + 1);
  hipLaunchKernelGGL(k_tridiagonal, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, fi,
                     lt, au, bu, cu, du, value, work, work + nk);
  return hipGetLastError();
}

}  // namespace gt
