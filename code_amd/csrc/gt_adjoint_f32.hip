// gt_adjoint_f32.hip -- the FP32 HALFSPLIT step one sub-step at a time, and
// its reverse mode (FP32 tracers and adjoints, FP64 layer heights).
//
// There is no counterpart in getm-model/code; the forward operator is the
// HALFSPLIT composition of advection_3d.F90:439-547 (u/2, v/2, w, v/2, u/2)
// with the sub-steps of adv_split_u.F90, adv_split_v.F90 and
// adv_split_w.F90, in the FP32 arithmetic of the batched strip kernels.
//
// Forward: one kernel per sub-step over all cells, built from the device
// functions of the strip kernels' FP32 flavour (face_setup, face_flux,
// cell_factors, cell_update_fast, wface_setup, wflux_fast in
// gt_strip_impl.h) with the same operands, so every intermediate state is the
// one the strip kernels hold at that point and the final state equals theirs
// bit for bit.  The layer heights D1..D5 depend on the velocities only and
// are formed once (k_heights), in FP64 with the sub-steps' expressions.
//
// Reverse: with velocities and layer heights held fixed each sub-step is a
// piecewise linear map of the tracer:
//   f_new(i) = f(i) + f(i)*am1(i) - (F(i) - F(i-1))*b(i)   (cell_update_fast)
//   F(X)     = vel(X)*rec(fuu, fu, fd) - kA(X)*(f(X+1) - f(X))
// with am1 = (Dio-Dn)/Dn and b = dt*splitfac*dlen*ard1/Dn.  Its transpose is
// gathered per cell (a thread owns a cell, or a column for w: no atomics).
// The limiter's derivative is that of the branch term the FP32 forward
// selected (the min or max of the terms times deltaf), one-sided at ties.
#include "gt_adjoint_f32.h"
#include "gt_strip_impl.h"

namespace gt {

// ---- layer heights ---------------------------------------------------------------
// D1..D5 from D0 = ho: adv_split_u.F90:275 (Di = Dio - dti*(U(i)*dy -
// U(i-1)*dy)*ard1), adv_split_v.F90 likewise with dx, adv_split_w.F90:288
// (hi = hio - dti*(ww(k) - ww(k-1))), where the sub-step updates the cell;
// level 0 keeps ho
__global__ void k_heights(Dims d, const uint8_t* __restrict__ flags, const double* __restrict__ ho,
                          const double* __restrict__ uu, const double* __restrict__ vv,
                          const double* __restrict__ ww, double* __restrict__ D1,
                          double* __restrict__ D2, double* __restrict__ D3,
                          double* __restrict__ D4, double* __restrict__ D5, double dt, double dx,
                          double dy, double ard1, int isub, int kshift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY || k > d.kmax) return;
  const long c2 = i + (long)d.NX * j, c = c2 + d.NXY * (long)k;
  const double D0 = ho[c];
  if (k == 0) {
    D1[c] = D0; D2[c] = D0; D3[c] = D0; D4[c] = D0; D5[c] = D0;
    return;
  }
  const int fl = flags[c2];
  const double dth = 0.5 * dt;
  const bool updu = (fl & MB_UUPD) && i >= 1 + isub && i <= d.NX - 2 - isub;
  const bool updv = (fl & MB_VUPD) && j >= 1 + isub && j <= d.NY - 2 - isub;
  const bool wupd = (fl & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2 && k <= d.kmax - kshift;
  const double du = updu ? uu[c] * dy - uu[c - 1] * dy : 0.0;
  const double dv = updv ? vv[c] * dx - vv[c - d.NX] * dx : 0.0;
  const double h1 = updu ? D0 - dth * du * ard1 : D0;
  const double h2 = updv ? h1 - dth * dv * ard1 : h1;
  const double h3 = wupd ? h2 - dt * (ww[c] - ww[c - d.NXY]) : h2;
  const double h4 = updv ? h3 - dth * dv * ard1 : h3;
  const double h5 = updu ? h4 - dth * du * ard1 : h4;
  D1[c] = h1; D2[c] = h2; D3[c] = h3; D4[c] = h4; D5[c] = h5;
}

// ---- a horizontal sub-step (FP32) ------------------------------------------------
// geometry of one sweep: dir 0 = u (faces along i, stride 1), 1 = v (NX)
struct Sweep {
  long st;       // element stride along the sweep
  int pos, n;    // position of the cell along the sweep, cells along it
  uint8_t mflux, mupd;
};
__device__ __forceinline__ Sweep sweep(const Dims& d, int dir, int i, int j) {
  return dir == 0 ? Sweep{1, i, d.NX, MB_UFLUX, MB_UUPD}
                  : Sweep{(long)d.NX, j, d.NY, MB_VFLUX, MB_VUPD};
}

// the constants of face X (between cells X and X+1 of the sweep) as the
// strip kernels form them; X must lie in 0..n-2
template <int SCH>
__device__ __forceinline__ FaceC<false> face_at(const Sweep& s, int X, long x2, long o,
                                                const uint8_t* flags, const double* vel,
                                                const double* Dv, const FaceK& k) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const bool active = X >= isub && X <= s.n - 2 - isub && (flags[x2] & s.mflux);
  const bool lim_l = SCH != UPSTREAM && X >= 1 && (flags[x2 - s.st] & s.mflux);
  const bool lim_r = SCH != UPSTREAM && (flags[x2 + s.st] & s.mflux);
  FaceC<false> fc;
  face_setup<SCH>(fc, active, lim_l, lim_r, vel[o + x2], Dv[o + x2], k);
  return fc;
}

// levels 0..kmax, one thread per cell; level 0 and cells the sub-step leaves
// alone are copied
template <int SCH, bool DIFF>
__global__ __launch_bounds__(256) void k_sub_h_f32(
    Dims d, int dir, int ntr, long ts, const float* __restrict__ fin, float* __restrict__ fout,
    const double* __restrict__ Dio, const double* __restrict__ Dn, const double* __restrict__ vel,
    const double* __restrict__ Dv, const uint8_t* __restrict__ flags, double dt,
    double dlen_flux, double dlen_face, double ard1, double AH) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY || k > d.kmax) return;
  const long c2 = i + (long)d.NX * j, o = d.NXY * (long)k, c = o + c2;
  const Sweep s = sweep(d, dir, i, j);
  const bool upd = k >= 1 && s.pos >= 1 + isub && s.pos <= s.n - 2 - isub && (flags[c2] & s.mupd);
  if (!upd) {
    for (int t = 0; t < ntr; t++) fout[t * ts + c] = fin[t * ts + c];
    return;
  }
  const FaceK fk = face_k(0.5 * dt, dlen_flux, AH);
  const FaceC<false> fcc = face_at<SCH>(s, s.pos, c2, o, flags, vel, Dv, fk);
  const FaceC<false> fcm = face_at<SCH>(s, s.pos - 1, c2 - s.st, o, flags, vel, Dv, fk);
  const CellK ck = cell_factors(Dio[c], Dn[c], dt * (0.5 * dlen_face * ard1));
  // cells pos-2..pos+2 (clamped into the array: UPSTREAM reaches past the
  // ends with operands it does not use)
  long q[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    int p = s.pos - 2 + m;
    p = p < 0 ? 0 : (p > s.n - 1 ? s.n - 1 : p);
    q[m] = c + (long)(p - s.pos) * s.st;
  }
  for (int t = 0; t < ntr; t++) {
    const float* f = fin + t * ts;
    const float fm2 = f[q[0]], fm1 = f[q[1]], f0 = f[q[2]], fp1 = f[q[3]], fp2 = f[q[4]];
    const float Fc = face_flux<float, SCH, DIFF>(fcc, fm1, f0, fp1, fp2, AH, dlen_flux);
    const float Fm = face_flux<float, SCH, DIFF>(fcm, fm2, fm1, f0, fp1, AH, dlen_flux);
    fout[t * ts + c] = cell_update_fast(ck, f0, Fc, Fm);
  }
}

// ---- the vertical sub-step (FP32) ------------------------------------------------
// at level K, as the strip kernels set them (adv_split_w.F90:249-300): the
// flux at K+1/2 exists below kmax, the one at K-1/2 above level 1 (or at
// level 1 with kshift); level K is updated up to kmax-kshift; the limiter
// applies above level 1-kshift going up and below kmax-1 going down
// tracers per thread of the forward kernel (the face constants are shared)
constexpr int kWGroup = 18;

template <int SCH>
__global__ __launch_bounds__(256) void k_sub_w_f32(Dims d, int ntr, long ts,
                                                   const float* __restrict__ fin,
                                                   float* __restrict__ fout,
                                                   const double* __restrict__ D2,
                                                   const double* __restrict__ D3,
                                                   const double* __restrict__ ww,
                                                   const uint8_t* __restrict__ flags, double dt,
                                                   int kshift) {
  // one thread per column and group of kWGroup tracers (blockIdx.z),
  // marching up; the face constants of a level serve the whole group
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int t0 = blockIdx.z * kWGroup;
  if (i >= d.NX || j >= d.NY || t0 >= ntr) return;
  const int t1 = min(t0 + kWGroup, ntr);
  const int kmax = d.kmax;
  const long NXY = d.NXY, c2 = i + (long)d.NX * j;
  const bool wcol = (flags[c2] & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2;
  if (!wcol) {
    for (int t = t0; t < t1; t++)
      for (int K = 0; K <= kmax; K++) fout[t * ts + c2 + NXY * K] = fin[t * ts + c2 + NXY * K];
    return;
  }
  for (int t = t0; t < t1; t++) fout[t * ts + c2] = fin[t * ts + c2];  // level 0
  double dm = D2[c2], dk = D2[c2 + NXY], dp = D2[c2 + 2 * NXY];  // D2 at K-1, K, K+1
  double wm = ww[c2], wkk = ww[c2 + NXY];                         // ww at K-1, K
  for (int K = 1; K <= kmax; K++) {
    const long cK = c2 + NXY * K;
    const bool wk_on = K <= kmax - 1, wp_on = K > 1 || kshift, kdo = K <= kmax - kshift;
    const bool limp_up = SCH != UPSTREAM && K - 1 > 1 - kshift;
    const bool limp_dn = SCH != UPSTREAM && K - 1 < kmax - 1;
    const bool limk_up = SCH != UPSTREAM && K > 1 - kshift;
    const bool limk_dn = SCH != UPSTREAM && K < kmax - 1;
    WFace<false> wfk, wfp;
    wface_setup<SCH>(wfk, wkk, dt, dk, K + 1 <= kmax ? dp : dk);
    wface_setup<SCH>(wfp, wm, dt, dm, dk);
    const CellK k3 = cell_factors(dk, D3[cK], dt);
    // the operands of the strip kernels: level -1 is zero, levels past kmax
    // are replaced by level K
    const long l0 = cK - 2 * NXY, l1 = cK - NXY;
    const long l3 = K + 1 <= kmax ? cK + NXY : cK, l4 = K + 2 <= kmax ? cK + 2 * NXY : cK;
    for (int t = t0; t < t1; t++) {
      const float* f = fin + t * ts;
      const float fKm2 = K > 1 ? f[l0] : 0.0f, fKm1 = f[l1], fK = f[cK];
      const float fKp1 = f[l3], fKp2 = f[l4];
      const float p = wflux_fast<SCH>(wfp, limp_up, limp_dn, fKm2, fKm1, fK, fKp1);
      const float q = wflux_fast<SCH>(wfk, limk_up, limk_dn, fKm1, fK, fKp1, fKp2);
      const float wprev = wp_on ? p : 0.0f, wk = wk_on ? q : 0.0f;
      fout[t * ts + cK] = kdo ? cell_update_fast(k3, fK, wk, wprev) : fK;
    }
    dm = dk; dk = dp;
    dp = K + 2 <= kmax ? D2[c2 + NXY * (K + 2)] : 0.0;
    wm = wkk;
    wkk = K + 1 <= kmax ? ww[c2 + NXY * (K + 1)] : 0.0;
  }
}

// ---- derivatives of the FP32 reconstruction -----------------------------------
struct RecGrad32 {
  float duu, du, dd;
};

// d rec_fast / d (fuu, fu, fd) at the forward point: the derivative of the
// term rec_fast selected (rec = fu + g*t, t one of the branch terms, each
// linear in du = fu-fuu and df = fd-fu), the identity where the limiter does
// not apply (g = 0) or deltaf*deltafu <= 0
template <int SCH>
__device__ __forceinline__ RecGrad32 rec_fast_grad(float fuu, float fu, float fd, float g, float A,
                                                   float al, float E) {
  if constexpr (SCH == UPSTREAM) {
    return {0.0f, 1.0f, 0.0f};
  } else {
    if (g == 0.0f) return {0.0f, 1.0f, 0.0f};
    const float df = fd - fu, du = fu - fuu;
    float tu, tf;  // d t / d du, d t / d df
    if constexpr (SCH == P2) {
      tu = 1.0f - al;
      tf = al;
    } else {
      if (!(df * du > 0.0f)) return {0.0f, 1.0f, 0.0f};
      const bool lo = df > 0.0f;  // the min of the terms (else the max)
      // pick(a, b): whether rec_fast's fminf (lo) / fmaxf (hi) returns b
      auto second = [&](float a, float b) { return lo ? b < a : b > a; };
      if constexpr (SCH == P2_PDM) {
        const float t1 = A * du, t2 = __builtin_fmaf(al, df - du, du), t3 = E * df;
        const bool s2 = second(t1, t2);
        const float m = s2 ? t2 : t1;
        if (second(m, t3)) { tu = 0.0f; tf = E; }
        else if (s2) { tu = 1.0f - al; tf = al; }
        else { tu = A; tf = 0.0f; }
      } else if constexpr (SCH == SUPERBEE) {
        // lo = max(min(2du, df), min(du, 2df)); hi = min(max(2du, df), max(du, 2df))
        const float a2 = 2.0f * du, b2 = 2.0f * df;
        const bool inner_a = lo ? (df < a2) : (df > a2);  // first pair: picks df
        const bool inner_b = lo ? (b2 < du) : (b2 > du);  // second pair: picks 2df
        const float va = inner_a ? df : a2, vb = inner_b ? b2 : du;
        const bool outer_b = lo ? (vb > va) : (vb < va);
        if (outer_b) { tu = inner_b ? 0.0f : 1.0f; tf = inner_b ? 2.0f : 0.0f; }
        else { tu = inner_a ? 0.0f : 2.0f; tf = inner_a ? 1.0f : 0.0f; }
      } else {  // MUSCL: lo = min(min(2du, h), 2df), hi = max(...), h = (df+du)/2
        const float a2 = 2.0f * du, b2 = 2.0f * df, h = 0.5f * (df + du);
        const bool s2 = second(a2, h);
        const float m = s2 ? h : a2;
        if (second(m, b2)) { tu = 0.0f; tf = 2.0f; }
        else if (s2) { tu = 0.5f; tf = 0.5f; }
        else { tu = 2.0f; tf = 0.0f; }
      }
    }
    // rec = fu + g*t(du, df), du = fu - fuu, df = fd - fu
    return {-g * tu, 1.0f + g * (tu - tf), g * tf};
  }
}

// partials of face_flux (fast flavour) with respect to the cells X-1..X+2
template <int SCH, bool DIFF>
__device__ __forceinline__ void face_flux_grad32(const FaceC<false>& fc, float fm1, float f0,
                                                 float fp1, float fp2, float out[4]) {
  const float fuu = fc.right ? fp2 : fm1;
  const float fu0 = fc.right ? fp1 : f0;
  const float fd = fc.right ? f0 : fp1;
  const RecGrad32 r = rec_fast_grad<SCH>(fuu, fu0, fd, fc.g, fc.A, fc.al, fc.E);
  if (fc.right) {
    out[0] = 0.0f; out[1] = fc.vel * r.dd; out[2] = fc.vel * r.du; out[3] = fc.vel * r.duu;
  } else {
    out[0] = fc.vel * r.duu; out[1] = fc.vel * r.du; out[2] = fc.vel * r.dd; out[3] = 0.0f;
  }
  if constexpr (DIFF) {
    out[1] += fc.kA;
    out[2] -= fc.kA;
  }
}

// ---- transpose of a horizontal sub-step ----------------------------------------
// f: the state before the sub-step; lam_in: adjoint of its output; lam_out:
// adjoint of its input.  Cell pos sits in the stencils of faces pos-2..pos+1.
template <int SCH, bool DIFF>
__global__ __launch_bounds__(256) void k_sub_h_f32_adj(
    Dims d, int dir, int ntr, long ts, const float* __restrict__ f,
    const float* __restrict__ lam_in, float* __restrict__ lam_out,
    const double* __restrict__ Dio, const double* __restrict__ Dn, const double* __restrict__ vel,
    const double* __restrict__ Dv, const uint8_t* __restrict__ flags, double dt,
    double dlen_flux, double dlen_face, double ard1, double AH) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= d.NX || j >= d.NY || k > d.kmax) return;
  const long c2 = i + (long)d.NX * j, o = d.NXY * (long)k, c = o + c2;
  if (k == 0) {
    for (int t = 0; t < ntr; t++) lam_out[t * ts + c] = lam_in[t * ts + c];
    return;
  }
  const Sweep s = sweep(d, dir, i, j);
  const FaceK fk = face_k(0.5 * dt, dlen_flux, AH);
  const double dtc = dt * (0.5 * dlen_face * ard1);
  // cells pos-2..pos+2: updated?, the factors of their update
  bool updm[5];
  float bm[5];
  float am1c = 0.0f;
  long q[7];  // element offsets of cells pos-3..pos+3 (clamped), q[m+1] = pos-2+m
#pragma unroll
  for (int m = 0; m < 7; m++) {
    int p = s.pos - 3 + m;
    p = p < 0 ? 0 : (p > s.n - 1 ? s.n - 1 : p);
    q[m] = c + (long)(p - s.pos) * s.st;
  }
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int p = s.pos - 2 + m;
    const long x2 = c2 + (long)(p - s.pos) * s.st;
    updm[m] = p >= 1 + isub && p <= s.n - 2 - isub && (flags[x2] & s.mupd);
    bm[m] = 0.0f;
    if (updm[m]) {
      const CellK ck = cell_factors(Dio[o + x2], Dn[o + x2], dtc);
      bm[m] = ck.b;
      if (m == 2) am1c = ck.am1;
    }
  }
  // faces X = pos-2..pos+1 (index m: X = pos-2+m); a face outside 0..n-2 or
  // inactive contributes nothing
  FaceC<false> fc[4];
  bool fon[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const int X = s.pos - 2 + m;
    fon[m] = X >= isub && X <= s.n - 2 - isub && (updm[m] || updm[m + 1]);
    if (fon[m]) {
      fc[m] = face_at<SCH>(s, X, c2 + (long)(X - s.pos) * s.st, o, flags, vel, Dv, fk);
    } else {
      fc[m] = FaceC<false>{};
    }
  }
  for (int t = 0; t < ntr; t++) {
    const float* ft = f + t * ts;
    const float* lt = lam_in + t * ts;
    const float lc = lt[c];
    float acc = updm[2] ? __builtin_fmaf(lc, am1c, lc) : lc;
    float fv[7];
#pragma unroll
    for (int m = 0; m < 7; m++) fv[m] = ft[q[m]];
    // mu(cell) = -b * lam: the adjoint of the flux through the cell's right
    // face (its left face enters with the opposite sign)
    float mu[5];
#pragma unroll
    for (int m = 0; m < 5; m++) mu[m] = updm[m] ? -bm[m] * lt[q[m + 1]] : 0.0f;
#pragma unroll
    for (int m = 0; m < 4; m++) {
      if (!fon[m]) continue;
      const float Fbar = mu[m] - mu[m + 1];
      float g[4];
      // face X = pos-2+m reads cells X-1..X+2 = fv[m..m+3]
      face_flux_grad32<SCH, DIFF>(fc[m], fv[m], fv[m + 1], fv[m + 2], fv[m + 3], g);
      // cell pos is stencil slot pos - X + 1 = 3 - m
      acc = __builtin_fmaf(Fbar, g[3 - m], acc);
    }
    lam_out[t * ts + c] = acc;
  }
}

// ---- transpose of the vertical sub-step ----------------------------------------
// one thread per column and tracer (blockIdx.z).  Marching over the faces
// K' = 0..kmax-1, the window holds levels K'-1..K'+2 (the flux at K'+1/2
// reads them): each level enters with its diagonal part (the adjoint through
// its own value) and mu = -b*lam (the adjoint of the flux through its upper
// face, into its own update), collects the face terms, and is stored when it
// leaves.  A face is the wk of level K' and the wprev of level K'+1 (K' >= 1,
// or K' = 0 with kshift).
template <int SCH>
__global__ __launch_bounds__(256) void k_sub_w_f32_adj(Dims d, int ntr, long ts,
                                                       const float* __restrict__ f,
                                                       const float* __restrict__ lam_in,
                                                       float* __restrict__ lam_out,
                                                       const double* __restrict__ D2,
                                                       const double* __restrict__ D3,
                                                       const double* __restrict__ ww,
                                                       const uint8_t* __restrict__ flags,
                                                       double dt, int kshift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int t = blockIdx.z;
  if (i >= d.NX || j >= d.NY || t >= ntr) return;
  const int kmax = d.kmax;
  const long NXY = d.NXY, c2 = i + (long)d.NX * j;
  const float* ft = f + t * ts;
  const float* lt = lam_in + t * ts;
  float* lo = lam_out + t * ts;
  const bool wcol = (flags[c2] & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2;
  if (!wcol) {
    for (int K = 0; K <= kmax; K++) lo[c2 + NXY * K] = lt[c2 + NXY * K];
    return;
  }
  struct Lev {
    float f, acc, mu;
  };
  auto enter = [&](int L) -> Lev {
    if (L > kmax) return {0.0f, 0.0f, 0.0f};
    const long cL = c2 + NXY * L;
    const float l = lt[cL];
    Lev v{ft[cL], l, 0.0f};
    if (L >= 1 && L <= kmax - kshift) {
      const CellK ck = cell_factors(D2[cL], D3[cL], dt);
      v.acc = __builtin_fmaf(l, ck.am1, l);
      v.mu = -ck.b * l;
    }
    return v;
  };
  Lev w0{0.0f, 0.0f, 0.0f}, w1 = enter(0), w2 = enter(1), w3 = enter(2);
  double d0 = D2[c2], d1 = D2[c2 + NXY];  // D2 at K', K'+1
  for (int Kf = 0; Kf <= kmax - 1; Kf++) {
    // level 0 is never updated (mu = 0), so at K' = 0 this is -mu(1) with
    // kshift and 0 without
    const float Fbar = (Kf >= 1 || kshift) ? w1.mu - w2.mu : 0.0f;
    if (Fbar != 0.0f) {
      WFace<false> wf;
      wface_setup<SCH>(wf, ww[c2 + NXY * Kf], dt, d0, d1);
      const bool lim_up = SCH != UPSTREAM && Kf > 1 - kshift;
      const bool lim_dn = SCH != UPSTREAM && Kf < kmax - 1;
      const bool up = wf.w > 0.0f;
      const float fkm1 = Kf >= 1 ? w0.f : 0.0f, fk = w1.f, fk1 = w2.f;
      const float fk2 = Kf + 2 <= kmax ? w3.f : fk;
      const float fu = up ? fk : fk1, fuu = up ? fkm1 : fk2, fd = up ? fk1 : fk;
      const bool lim = up ? lim_up : lim_dn;
      const RecGrad32 r = rec_fast_grad<SCH>(fuu, fu, fd, lim ? wf.g : 0.0f, wf.A, wf.al, wf.E);
      const float a = Fbar * wf.w;
      // (a slot past either end of the column is never stored)
      if (up) {
        w0.acc = __builtin_fmaf(a, r.duu, w0.acc);
        w1.acc = __builtin_fmaf(a, r.du, w1.acc);
        w2.acc = __builtin_fmaf(a, r.dd, w2.acc);
      } else {
        w1.acc = __builtin_fmaf(a, r.dd, w1.acc);
        w2.acc = __builtin_fmaf(a, r.du, w2.acc);
        w3.acc = __builtin_fmaf(a, r.duu, w3.acc);
      }
    }
    if (Kf >= 1) lo[c2 + NXY * (Kf - 1)] = w0.acc;
    w0 = w1; w1 = w2; w2 = w3;
    w3 = enter(Kf + 3);
    d0 = d1;
    d1 = Kf + 2 <= kmax ? D2[c2 + NXY * (Kf + 2)] : 0.0;
  }
  lo[c2 + NXY * (kmax - 1)] = w0.acc;
  lo[c2 + NXY * kmax] = w1.acc;
}

// ---- launchers --------------------------------------------------------------------
static inline dim3 grid3(const Dims& d, int nz) {
  return dim3((d.NX + 63) / 64, (d.NY + 3) / 4, nz);
}

hipError_t launch_heights(hipStream_t st, Dims d, const uint8_t* flags, const double* ho,
                          const double* uu, const double* vv, const double* ww,
                          double* const D[5], double dt, double dx, double dy, double ard1,
                          int scheme, int kshift) {
  hipLaunchKernelGGL(k_heights, grid3(d, d.kmax + 1), dim3(64, 4), 0, st, d, flags, ho, uu, vv,
                     ww, D[0], D[1], D[2], D[3], D[4], dt, dx, dy, ard1,
                     scheme == UPSTREAM ? 0 : 1, kshift);
  return hipGetLastError();
}

template <bool ADJ, int SCH, bool DIFF>
static void launch_h_one(hipStream_t st, Dims d, int dir, int ntr, long ts, const float* a,
                         const float* b, float* out, const double* Dio, const double* Dn,
                         const double* vel, const double* Dv, const uint8_t* flags, double dt,
                         double lf, double lc, double ard1, double AH) {
  if constexpr (ADJ) {
    hipLaunchKernelGGL((k_sub_h_f32_adj<SCH, DIFF>), grid3(d, d.kmax + 1), dim3(64, 4), 0, st, d,
                       dir, ntr, ts, a, b, out, Dio, Dn, vel, Dv, flags, dt, lf, lc, ard1, AH);
  } else {
    (void)b;
    hipLaunchKernelGGL((k_sub_h_f32<SCH, DIFF>), grid3(d, d.kmax + 1), dim3(64, 4), 0, st, d,
                       dir, ntr, ts, a, out, Dio, Dn, vel, Dv, flags, dt, lf, lc, ard1, AH);
  }
}

template <bool ADJ>
static hipError_t launch_h(hipStream_t st, Dims d, int dir, int ntr, long ts, const float* a,
                           const float* b, float* out, const double* Dio, const double* Dn,
                           const double* vel, const double* Dv, const uint8_t* flags, double dt,
                           double dx, double dy, double ard1, int scheme, double AH) {
  const double lf = dir == 0 ? dx : dy, lc = dir == 0 ? dy : dx;
#define GT_H(S)                                                                              \
  (AH > 0.0 ? launch_h_one<ADJ, S, true>(st, d, dir, ntr, ts, a, b, out, Dio, Dn, vel, Dv,    \
                                         flags, dt, lf, lc, ard1, AH)                        \
            : launch_h_one<ADJ, S, false>(st, d, dir, ntr, ts, a, b, out, Dio, Dn, vel, Dv,   \
                                          flags, dt, lf, lc, ard1, AH))
  switch (scheme) {
    case UPSTREAM: GT_H(UPSTREAM); break;
    case P2: GT_H(P2); break;
    case SUPERBEE: GT_H(SUPERBEE); break;
    case MUSCL: GT_H(MUSCL); break;
    case P2_PDM: GT_H(P2_PDM); break;
    default: return hipErrorInvalidValue;
  }
#undef GT_H
  return hipGetLastError();
}

hipError_t launch_sub_h_f32(hipStream_t st, Dims d, int dir, int ntr, long ts, const float* fin,
                            float* fout, const double* Dio, const double* Dn, const double* vel,
                            const double* Dv, const uint8_t* flags, double dt, double dx,
                            double dy, double ard1, int scheme, double AH) {
  return launch_h<false>(st, d, dir, ntr, ts, fin, nullptr, fout, Dio, Dn, vel, Dv, flags, dt, dx,
                         dy, ard1, scheme, AH);
}

hipError_t launch_sub_h_f32_adj(hipStream_t st, Dims d, int dir, int ntr, long ts,
                                const float* f, const float* lam_in, float* lam_out,
                                const double* Dio, const double* Dn, const double* vel,
                                const double* Dv, const uint8_t* flags, double dt, double dx,
                                double dy, double ard1, int scheme, double AH) {
  return launch_h<true>(st, d, dir, ntr, ts, f, lam_in, lam_out, Dio, Dn, vel, Dv, flags, dt, dx,
                        dy, ard1, scheme, AH);
}

template <bool ADJ, int SCH>
static void launch_w_one(hipStream_t st, Dims d, int ntr, long ts, const float* a,
                         const float* b, float* out, const double* D2, const double* D3,
                         const double* ww, const uint8_t* flags, double dt, int kshift) {
  if constexpr (ADJ) {
    hipLaunchKernelGGL((k_sub_w_f32_adj<SCH>), grid3(d, ntr), dim3(64, 4), 0, st, d, ntr, ts, a, b,
                       out, D2, D3, ww, flags, dt, kshift);
  } else {
    (void)b;
    hipLaunchKernelGGL((k_sub_w_f32<SCH>), grid3(d, (ntr + kWGroup - 1) / kWGroup), dim3(64, 4),
                       0, st, d, ntr, ts, a, out,
                       D2, D3, ww, flags, dt, kshift);
  }
}

template <bool ADJ>
static hipError_t launch_w(hipStream_t st, Dims d, int ntr, long ts, const float* a,
                           const float* b, float* out, const double* D2, const double* D3,
                           const double* ww, const uint8_t* flags, double dt, int scheme,
                           int kshift) {
  switch (scheme) {
    case UPSTREAM: launch_w_one<ADJ, UPSTREAM>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case P2: launch_w_one<ADJ, P2>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case SUPERBEE: launch_w_one<ADJ, SUPERBEE>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case MUSCL: launch_w_one<ADJ, MUSCL>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case P2_PDM: launch_w_one<ADJ, P2_PDM>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_sub_w_f32(hipStream_t st, Dims d, int ntr, long ts, const float* fin,
                            float* fout, const double* D2, const double* D3, const double* ww,
                            const uint8_t* flags, double dt, int scheme, int kshift) {
  return launch_w<false>(st, d, ntr, ts, fin, nullptr, fout, D2, D3, ww, flags, dt, scheme,
                         kshift);
}

hipError_t launch_sub_w_f32_adj(hipStream_t st, Dims d, int ntr, long ts, const float* f,
                                const float* lam_in, float* lam_out, const double* D2,
                                const double* D3, const double* ww, const uint8_t* flags,
                                double dt, int scheme, int kshift) {
  return launch_w<true>(st, d, ntr, ts, f, lam_in, lam_out, D2, D3, ww, flags, dt, scheme,
                        kshift);
}

}  // namespace gt
