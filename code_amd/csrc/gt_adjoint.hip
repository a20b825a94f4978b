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

// ---- Shchepetkin-McWilliams (ip_shchepetkin_mcwilliams.F90:90-286) ------------
// The pressure gradient is nonlinear in the buoyancy through the harmonic
// means of its differences (zero where 2ab <= eps); the transpose is that of
// its linearisation at the forward buoyancy (layer heights, zz, hun, hvn and
// sseo fixed).  The reference's default-REAL literals, as the forward uses
// them:
#define SMA_EPS ((double)1.e-10f)
#define SMA_ONEFIFTH ((double)0.2f)
#define SMA_C12 ((double)(1.0f / 12.0f))

// harmonic mean 2ac/(a+c) with the eps test, and its partials
struct HmG {
  double da, dc;
};
__device__ __forceinline__ HmG sm_hm_grad(double a, double c) {
  if (!(2.0 * a * c > SMA_EPS)) return {0.0, 0.0};
  const double s = 1.0 / (a + c);
  return {2.0 * c * c * s * s, 2.0 * a * a * s * s};
}
__device__ __forceinline__ double sm_hm_x(double a, double c) {  // the forward's, with eps
  const double cff = 2.0 * a * c;
  return cff > SMA_EPS ? cff * (1.0 / (a + c)) : 0.0;
}

// d(FCbar*FC)/d b at the four cells X-1..X+2 of a face (FC of :153-211),
// FCbar the adjoint of the face's AJ
__device__ __forceinline__ void sm_face_grad(const double b[4], const double z[4], double FCbar,
                                             double out[4]) {
  const double zrm = z[1] - z[0], zr0 = z[2] - z[1], zrp = z[3] - z[2];
  const double dZx0 = sm_hm_x(zrm, zr0), dZx1 = sm_hm_x(zr0, zrp);
  const double Bz = zr0 - SMA_C12 * (dZx1 + dZx0), Cz = dZx1 - dZx0;
  const double a0 = b[1] - b[0], br0 = b[2] - b[1], c1 = b[3] - b[2];
  const HmG g0 = sm_hm_grad(a0, br0), g1 = sm_hm_grad(br0, c1);
  const double R0 = FCbar * 0.5 * SMA_ONEFIFTH * (Bz - SMA_C12 * Cz);   // d/d dRx0
  const double R1 = -FCbar * 0.5 * SMA_ONEFIFTH * (Bz + SMA_C12 * Cz);  // d/d dRx1
  const double a0b = R0 * g0.da, c0b = R0 * g0.dc, a1b = R1 * g1.da, c1b = R1 * g1.dc;
  const double brb = c0b + a1b + FCbar * 0.5 * SMA_ONEFIFTH * Cz;
  out[0] = -a0b;
  out[1] = FCbar * 0.5 * zr0 - brb + a0b;
  out[2] = FCbar * 0.5 * zr0 + brb - c1b;
  out[3] = c1b;
}

// horizontal part, one thread per cell and level: bdir receives the
// buoyancy adjoint through FC (the faces i-2..i+1 and j-2..j+1 whose stencil
// holds the cell), Pbar the adjoint of P(i,j,k) (the faces i-1, i, j-1, j)
__global__ void k_sm_adj_h(Dims d, const int32_t* __restrict__ au, const int32_t* __restrict__ av,
                           const double* __restrict__ buoy, const double* __restrict__ zz,
                           const double* __restrict__ hun, const double* __restrict__ hvn,
                           const double* __restrict__ ixbar, const double* __restrict__ iybar,
                           double* __restrict__ bdir, double* __restrict__ Pbar, double dxm1,
                           double dym1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY || k > d.kmax) return;
  const long NX = d.NX, o = d.NXY * (long)k, c = o + i + NX * j;
  if (k == 0) {
    bdir[c] = 0.0;
    Pbar[c] = 0.0;
    return;
  }
  double acc = 0.0, pb = 0.0;
  // a face is computed where the forward computes it: x in 2..NX-3 and y in
  // 2..NY-3 at its own cell, the mask >= 1
  auto inner = [&](int x, int y) { return x >= 2 && x <= d.NX - 3 && y >= 2 && y <= d.NY - 3; };
  for (int dir = 0; dir < 2; dir++) {
    const long st = dir == 0 ? 1 : NX;
    const int pos = dir == 0 ? i : j;
    const int n = dir == 0 ? d.NX : d.NY;
    const int32_t* m = dir == 0 ? au : av;
    const double* h = dir == 0 ? hun : hvn;
    const double* ub = dir == 0 ? ixbar : iybar;
    const double dm1 = dir == 0 ? dxm1 : dym1;
    for (int X = pos - 2; X <= pos + 1; X++) {
      if (X < 2 || X > n - 3) continue;
      const int fx = dir == 0 ? X : i, fy = dir == 0 ? j : X;
      if (!inner(fx, fy)) continue;
      const long f2 = fx + NX * (long)fy;
      if (m[f2] < 1) continue;
      const double AJb = ub[o + f2] * dm1 * h[o + f2];
      if (AJb == 0.0) continue;
      double b[4], z[4], g[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long cq = o + f2 + (long)(q - 1) * st;
        b[q] = buoy[cq];
        z[q] = zz[cq];
      }
      sm_face_grad(b, z, AJb, g);
      acc += g[pos - X + 1];
      if (X == pos) pb -= AJb;      // AJ = P(X+1) - P(X)
      if (X == pos - 1) pb += AJb;
    }
  }
  bdir[c] = acc;
  Pbar[c] = pb;
}

// column part, one thread per column: the transpose of the downward P
// recurrence (:108-146) with its vertical harmonic means, plus the surface
// term -h dm1 (sseo(+) - sseo) 0.5 (b(+,kmax) + b(kmax)) of every level's
// faces; bbar = bdir + these (kmax <= 64)
__global__ void k_sm_adj_col(Dims d, const int32_t* __restrict__ az,
                             const int32_t* __restrict__ au, const int32_t* __restrict__ av,
                             const double* __restrict__ buoy, const double* __restrict__ zz,
                             const double* __restrict__ hn, const double* __restrict__ hun,
                             const double* __restrict__ hvn, const double* __restrict__ sseo,
                             const double* __restrict__ ixbar, const double* __restrict__ iybar,
                             const double* __restrict__ bdir, const double* __restrict__ Pbar,
                             double* __restrict__ bbar, double dxm1, double dym1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= d.NX || j >= d.NY) return;
  const long NX = d.NX, NXY = d.NXY, c0 = i + NX * j;
  const int kmax = d.kmax;
  double acc[64 + 1];
  for (int k = 0; k <= kmax; k++) acc[k] = bdir[c0 + NXY * k];
  auto inner = [&](int x, int y) { return x >= 2 && x <= d.NX - 3 && y >= 2 && y <= d.NY - 3; };
  // surface term: this column is the + side of face i-1 (j-1) and the own
  // side of face i (j)
  auto surf = [&](long f2, const int32_t* m, const double* h, const double* ub, long st,
                  double dm1) {
    if (m[f2] < 1) return;
    double sum = 0.0;
    for (int k = 1; k <= kmax; k++) sum += ub[f2 + NXY * k] * h[f2 + NXY * k];
    acc[kmax] += -dm1 * (sseo[f2 + st] - sseo[f2]) * 0.5 * sum;
  };
  if (inner(i, j)) surf(c0, au, hun, ixbar, 1, dxm1);
  if (i >= 1 && inner(i - 1, j)) surf(c0 - 1, au, hun, ixbar, 1, dxm1);
  if (inner(i, j)) surf(c0, av, hvn, iybar, NX, dym1);
  if (j >= 1 && inner(i, j - 1)) surf(c0 - NX, av, hvn, iybar, NX, dym1);
  // the vertical part where the forward forms P
  if (i >= 2 && i <= d.NX - 2 && j >= 2 && j <= d.NY - 2 && az[c0] >= 1) {
#define AT(a, k) (a)[c0 + NXY * (long)(k)]
    double db[64 + 1];   // adjoint of the raw differences d_k = b(k+1) - b(k)
    double Hb[64 + 1];   // adjoint of the harmonic means H_k
    for (int k = 0; k <= kmax; k++) db[k] = Hb[k] = 0.0;
    auto dR = [&](int k) -> double {  // d_k with d_kmax = d_kmax-1, d_0 = d_1
      const int kk = k >= kmax ? kmax - 1 : (k <= 0 ? 1 : k);
      return AT(buoy, kk + 1) - AT(buoy, kk);
    };
    auto dZ = [&](int k) -> double {
      const int kk = k >= kmax ? kmax - 1 : (k <= 0 ? 1 : k);
      return AT(zz, kk + 1) - AT(zz, kk);
    };
    auto hZ = [&](int k) -> double {  // H of the layer-centre differences (no eps)
      const double a = dZ(k), b = dZ(k - 1);
      return 2.0 * a * b / (a + b);
    };
    // C_m: the adjoint of G_m = P(m) - P(m+1), the sum of Pbar(k) for k <= m
    double C = 0.0;
    for (int m = 1; m <= kmax - 1; m++) {
      C += AT(Pbar, m);
      const double zr = AT(zz, m + 1) - AT(zz, m);
      const double hz1 = hZ(m + 1), hz0 = hZ(m);
      const double Bz = zr - SMA_C12 * (hz1 + hz0), Dz = hz1 - hz0;
      acc[m + 1] += C * 0.5 * zr;
      acc[m] += C * 0.5 * zr;
      db[m] += C * 0.5 * SMA_ONEFIFTH * Dz;
      Hb[m + 1] += -C * 0.5 * SMA_ONEFIFTH * (Bz + SMA_C12 * Dz);
      Hb[m] += C * 0.5 * SMA_ONEFIFTH * (Bz - SMA_C12 * Dz);
    }
    C += AT(Pbar, kmax);
    // P(kmax) = (b(kmax) + 0.25 ht (b(kmax)-b(kmax-1))/(zz(kmax)-zz(kmax-1))) 0.5 ht
    {
      const double ht = AT(hn, kmax);
      const double q = 0.5 * 0.5 * ht * (1.0 / (AT(zz, kmax) - AT(zz, kmax - 1)));
      acc[kmax] += C * 0.5 * ht * (1.0 + q);
      acc[kmax - 1] += -C * 0.5 * ht * q;
    }
    // H_k = hm(d_k, d_k-1): into the raw differences, the copies d_kmax and
    // d_0 folded into d_kmax-1 and d_1
    for (int k = 1; k <= kmax; k++) {
      if (Hb[k] == 0.0) continue;
      const HmG g = sm_hm_grad(dR(k), dR(k - 1));
      const int ka = k >= kmax ? kmax - 1 : k, kc = k - 1 <= 0 ? 1 : k - 1;
      db[ka] += Hb[k] * g.da;
      db[kc] += Hb[k] * g.dc;
    }
    for (int k = 1; k <= kmax - 1; k++) {
      acc[k + 1] += db[k];
      acc[k] -= db[k];
    }
#undef AT
  }
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

hipError_t launch_sm_adj(hipStream_t st, Dims d, const int32_t* az, const int32_t* au,
                         const int32_t* av, const double* buoy, const double* zz, const double* hn,
                         const double* hun, const double* hvn, const double* sseo,
                         const double* ixbar, const double* iybar, double* bdir, double* Pbar,
                         double* bbar, double dx, double dy) {
  if (d.kmax > 64 || d.kmax < 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_sm_adj_h, g2(d, d.kmax + 1), dim3(64, 4), 0, st, d, au, av, buoy, zz, hun,
                     hvn, ixbar, iybar, bdir, Pbar, 1.0 / dx, 1.0 / dy);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sm_adj_col, g2(d, 1), dim3(64, 4), 0, st, d, az, au, av, buoy, zz, hn,
                     hun, hvn, sseo, ixbar, iybar, bdir, Pbar, bbar, 1.0 / dx, 1.0 / dy);
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
