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
// formed face by face, every face once: along i a lane owns a face and the
// cells collect their faces' terms by cross-lane moves; along j and k a
// thread marches rows or levels with a four-deep window in registers and
// stores each output when no further face reaches it (no atomics).  The
// limiter's derivative is that of the branch term the FP32 forward selected
// (the min or max of the terms times deltaf), one-sided at ties.
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

// the u half sub-step: a wavefront holds 64 consecutive cells of a row, lane L
// forms the flux of face x from the cells x-1..x+2 (cross-lane moves) and
// cell x takes the flux of face x-1 from lane L-1, as the strip kernels do:
// lanes 2..61 are complete, 60 cells per wavefront.  Level 0 and cells the
// sub-step leaves alone are copied.
constexpr int kUFwdOut = 60;

template <int SCH, bool DIFF>
__global__ __launch_bounds__(256) void k_sub_u_f32(
    Dims d, int ntr, long ts, const float* __restrict__ fin, float* __restrict__ fout,
    const double* __restrict__ Dio, const double* __restrict__ Dn, const double* __restrict__ uu,
    const double* __restrict__ hu, const uint8_t* __restrict__ flags, double dt, double dx,
    double dy, double ard1, double AH) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const int lane = threadIdx.x;  // blockDim (64, 4): one row per wavefront
  const int x = blockIdx.x * kUFwdOut - 2 + lane;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (j >= d.NY) return;  // the whole wavefront
  const bool xin = x >= 0 && x < d.NX;
  const long c2 = (long)min(max(x, 0), d.NX - 1) + (long)d.NX * j;
  const long o = d.NXY * (long)k, c = o + c2;
  const bool out = xin && lane >= 2 && lane <= 61;
  if (k == 0) {
    if (out)
      for (int t = 0; t < ntr; t++) fout[t * ts + c] = fin[t * ts + c];
    return;
  }
  const Sweep s = sweep(d, 0, x, j);
  FaceC<false> fc{};
  if (xin && x >= isub && x <= d.NX - 2 - isub)
    fc = face_at<SCH>(s, x, c2, o, flags, uu, hu, face_k(0.5 * dt, dx, AH));
  const bool upd = xin && x >= 1 + isub && x <= d.NX - 2 - isub && (flags[c2] & MB_UUPD);
  CellK ck{0.0f, 0.0f};
  if (upd) ck = cell_factors(Dio[c], Dn[c], dt * (0.5 * dy * ard1));
  for (int t = 0; t < ntr; t++) {
    const float fx = fin[t * ts + c];
    const float fm1 = from_lower_lane(fx), fp1 = from_upper_lane(fx);
    const float fp2 = from_upper_lane(fp1);
    const float Fc = face_flux<float, SCH, DIFF>(fc, fm1, fx, fp1, fp2, AH, dx);
    const float Fm = from_lower_lane(Fc);
    const float u = cell_update_fast(ck, fx, Fc, Fm);
    if (out) fout[t * ts + c] = upd ? u : fx;
  }
}

// the v half sub-step: a wavefront holds 64 columns and marches the faces
// Y = y0-1 .. y1-1 for a group of G tracers, with the rows Y-1..Y+2 and the
// flux of face Y-1 in registers; row Y is updated after face Y
constexpr int kVRows = 16;
constexpr int kVGroup = 6;

template <int SCH, bool DIFF>
__global__ __launch_bounds__(64) void k_sub_v_f32(
    Dims d, int ntr, long ts, const float* __restrict__ fin, float* __restrict__ fout,
    const double* __restrict__ Dio, const double* __restrict__ Dn, const double* __restrict__ vv,
    const double* __restrict__ hv, const uint8_t* __restrict__ flags, double dt, double dx,
    double dy, double ard1, double AH, int ngroups) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  constexpr int G = kVGroup;
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y0 = blockIdx.y * kVRows;
  const int k = blockIdx.z / ngroups, t0 = (blockIdx.z % ngroups) * G;
  if (x >= d.NX) return;  // no cross-lane moves here
  const int NX = d.NX, NY = d.NY;
  const long o = d.NXY * (long)k;
  const int y1 = min(y0 + kVRows, NY);
  long tb[G];
#pragma unroll
  for (int g = 0; g < G; g++) tb[g] = (long)min(t0 + g, ntr - 1) * ts;
  if (k == 0) {
    for (int y = y0; y < y1; y++)
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) fout[tb[g] + x + (long)NX * y] = fin[tb[g] + x + (long)NX * y];
    return;
  }
  const Sweep s{(long)NX, 0, NY, MB_VFLUX, MB_VUPD};
  const FaceK fk = face_k(0.5 * dt, dy, AH);
  const double dtc = dt * (0.5 * dx * ard1);
  auto row = [&](int y) -> long { return o + x + (long)NX * (y >= 0 && y < NY ? y : 0); };
  float w[G][4], Fp[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    Fp[g] = 0.0f;
#pragma unroll
    for (int q = 0; q < 4; q++) w[g][q] = fin[tb[g] + row(y0 - 2 + q)];
  }
  for (int Y = y0 - 1; Y < y1; Y++) {
    FaceC<false> fc{};
    if (Y >= isub && Y <= NY - 2 - isub) fc = face_at<SCH>(s, Y, x + (long)NX * Y, o, flags, vv, hv, fk);
    const bool st = Y >= y0;
    const long cY = o + x + (long)NX * (Y >= 0 ? Y : 0);
    const bool upd = st && Y >= 1 + isub && Y <= NY - 2 - isub && (flags[cY - o] & MB_VUPD);
    CellK ck{0.0f, 0.0f};
    if (upd) ck = cell_factors(Dio[cY], Dn[cY], dtc);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const float F = face_flux<float, SCH, DIFF>(fc, w[g][0], w[g][1], w[g][2], w[g][3], AH, dy);
      if (st && t0 + g < ntr) {
        const float u = cell_update_fast(ck, w[g][1], F, Fp[g]);
        fout[tb[g] + cY] = upd ? u : w[g][1];
      }
      Fp[g] = F;
      w[g][0] = w[g][1];
      w[g][1] = w[g][2];
      w[g][2] = w[g][3];
      w[g][3] = fin[tb[g] + row(Y + 3)];
    }
  }
}

// ---- the vertical sub-step (FP32) ------------------------------------------------
// at level K, as the strip kernels set them (adv_split_w.F90:249-300): the
// flux at K+1/2 exists below kmax, the one at K-1/2 above level 1 (or at
// level 1 with kshift); level K is updated up to kmax-kshift; the limiter
// applies above level 1-kshift going up and below kmax-1 going down
template <int SCH>
__global__ __launch_bounds__(256) void k_sub_w_f32(Dims d, int ntr, long ts,
                                                   const float* __restrict__ fin,
                                                   float* __restrict__ fout,
                                                   const double* __restrict__ D2,
                                                   const double* __restrict__ D3,
                                                   const double* __restrict__ ww,
                                                   const uint8_t* __restrict__ flags, double dt,
                                                   int kshift) {
  // one thread per column and group of kVGroup tracers (blockIdx.z), marching
  // up with the levels K-2..K+2 of each tracer, K-1..K+1 of D2 and K-1, K of
  // ww in registers (kmax >= 2); the face constants serve the whole group
  constexpr int G = kVGroup;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int t0 = blockIdx.z * G;
  if (i >= d.NX || j >= d.NY || t0 >= ntr) return;
  const int kmax = d.kmax;
  const long NXY = d.NXY, c2 = i + (long)d.NX * j;
  long tb[G];
#pragma unroll
  for (int g = 0; g < G; g++) tb[g] = (long)min(t0 + g, ntr - 1) * ts;
  const bool wcol = (flags[c2] & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2;
  if (!wcol) {
    for (int K = 0; K <= kmax; K++)
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) fout[tb[g] + c2 + NXY * K] = fin[tb[g] + c2 + NXY * K];
    return;
  }
  float n[G][5];  // levels K-2..K+2 (entries past kmax unused)
#pragma unroll
  for (int g = 0; g < G; g++) {
    const float* f = fin + tb[g];
    n[g][0] = 0.0f;
    n[g][1] = f[c2];
    n[g][2] = f[c2 + NXY];
    n[g][3] = f[c2 + 2 * NXY];
    n[g][4] = kmax >= 3 ? f[c2 + 3 * NXY] : 0.0f;
    if (t0 + g < ntr) fout[tb[g] + c2] = n[g][1];  // level 0
  }
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
    const bool nxt = K + 3 <= kmax;
    const long cn = c2 + NXY * (nxt ? K + 3 : kmax);
#pragma unroll
    for (int g = 0; g < G; g++) {
      // the operands of the strip kernels: level -1 is zero, levels past
      // kmax are replaced by level K
      const float fKm2 = K > 1 ? n[g][0] : 0.0f, fKm1 = n[g][1], fK = n[g][2];
      const float fKp1 = K + 1 <= kmax ? n[g][3] : fK, fKp2 = K + 2 <= kmax ? n[g][4] : fK;
      const float p = wflux_fast<SCH>(wfp, limp_up, limp_dn, fKm2, fKm1, fK, fKp1);
      const float q = wflux_fast<SCH>(wfk, limk_up, limk_dn, fKm1, fK, fKp1, fKp2);
      const float wprev = wp_on ? p : 0.0f, wk = wk_on ? q : 0.0f;
      const float r = kdo ? cell_update_fast(k3, fK, wk, wprev) : fK;
      if (t0 + g < ntr) fout[tb[g] + cK] = r;
      n[g][0] = n[g][1];
      n[g][1] = n[g][2];
      n[g][2] = n[g][3];
      n[g][3] = n[g][4];
      n[g][4] = nxt ? fin[tb[g] + cn] : 0.0f;
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

// ---- transpose of the u sub-step: one face per lane, cross-lane sums --------
// A wavefront holds 64 consecutive cells of a row; lane L owns cell x and the
// face x (between x and x+1).  Face x reads the cells x-1..x+2 and its flux
// enters the updates of cells x and x+1; cell x collects the faces x-2..x+1
// (lanes L-2..L+1).  So lanes 3..60 are complete: 58 cells per wavefront.
constexpr int kUOut = 58;

template <int SCH, bool DIFF>
__global__ __launch_bounds__(256) void k_sub_u_f32_adj(
    Dims d, int ntr, long ts, const float* __restrict__ f, const float* __restrict__ lam_in,
    float* __restrict__ lam_out, const double* __restrict__ Dio, const double* __restrict__ Dn,
    const double* __restrict__ uu, const double* __restrict__ hu,
    const uint8_t* __restrict__ flags, double dt, double dx, double dy, double ard1, double AH) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const int lane = threadIdx.x;  // blockDim (64, 4): one row per wavefront
  const int x = blockIdx.x * kUOut - 3 + lane;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (j >= d.NY) return;  // the whole wavefront
  const bool xin = x >= 0 && x < d.NX;
  const long c2 = (long)min(max(x, 0), d.NX - 1) + (long)d.NX * j;
  const long o = d.NXY * (long)k, c = o + c2;
  const bool out = xin && lane >= 3 && lane <= 60;
  if (k == 0) {
    if (out)
      for (int t = 0; t < ntr; t++) lam_out[t * ts + c] = lam_in[t * ts + c];
    return;
  }
  const Sweep s = sweep(d, 0, x, j);
  FaceC<false> fc{};
  if (xin && x >= isub && x <= d.NX - 2 - isub)
    fc = face_at<SCH>(s, x, c2, o, flags, uu, hu, face_k(0.5 * dt, dx, AH));
  const bool upd = xin && x >= 1 + isub && x <= d.NX - 2 - isub && (flags[c2] & MB_UUPD);
  CellK ck{0.0f, 0.0f};
  if (upd) ck = cell_factors(Dio[c], Dn[c], dt * (0.5 * dy * ard1));
  for (int t = 0; t < ntr; t++) {
    const float fx = f[t * ts + c], lx = lam_in[t * ts + c];
    const float fm1 = from_lower_lane(fx), fp1 = from_upper_lane(fx);
    const float fp2 = from_upper_lane(fp1);
    const float mu = upd ? -ck.b * lx : 0.0f;  // adjoint of the flux through face x into cell x
    const float Fbar = mu - from_upper_lane(mu);
    float g[4];
    face_flux_grad32<SCH, DIFF>(fc, fm1, fx, fp1, fp2, g);
    float acc = upd ? __builtin_fmaf(lx, ck.am1, lx) : lx;
    acc = __builtin_fmaf(Fbar, g[1], acc);                      // face x
    acc += from_upper_lane(Fbar * g[0]);                         // face x+1
    acc += from_lower_lane(Fbar * g[2]);                         // face x-1
    acc += from_lower_lane(from_lower_lane(Fbar * g[3]));        // face x-2
    if (out) lam_out[t * ts + c] = acc;
  }
}

// ---- transpose of the v sub-step: a row march -----------------------------------
// A wavefront holds 64 columns and marches the faces Y = y0-2 .. y0+R (the
// faces of the output rows y0..y0+R-1) for a group of G tracers.  Before face
// Y the window holds rows Y-1..Y+2: each row enters with its diagonal part
// and mu = -b*lam, collects the terms of the faces whose stencil holds it, and
// is stored when it leaves (row Y-1 after face Y).
template <int SCH, bool DIFF>
__global__ __launch_bounds__(64) void k_sub_v_f32_adj(
    Dims d, int ntr, long ts, const float* __restrict__ f, const float* __restrict__ lam_in,
    float* __restrict__ lam_out, const double* __restrict__ Dio, const double* __restrict__ Dn,
    const double* __restrict__ vv, const double* __restrict__ hv,
    const uint8_t* __restrict__ flags, double dt, double dx, double dy, double ard1, double AH,
    int ngroups) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  constexpr int G = kVGroup;
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y0 = blockIdx.y * kVRows;
  const int k = blockIdx.z / ngroups, t0 = (blockIdx.z % ngroups) * G;
  if (x >= d.NX) return;  // no cross-lane moves here
  const int NX = d.NX, NY = d.NY;
  const long o = d.NXY * (long)k;
  const int y1 = min(y0 + kVRows, NY);  // output rows y0..y1-1
  long tb[G];
#pragma unroll
  for (int g = 0; g < G; g++) tb[g] = (long)min(t0 + g, ntr - 1) * ts;
  if (k == 0) {
    for (int y = y0; y < y1; y++)
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) lam_out[tb[g] + x + (long)NX * y] = lam_in[tb[g] + x + (long)NX * y];
    return;
  }
  const Sweep s{(long)NX, 0, NY, MB_VFLUX, MB_VUPD};
  const FaceK fk = face_k(0.5 * dt, dy, AH);
  const double dtc = dt * (0.5 * dx * ard1);
  struct Row {
    float f, acc, mu;
  };
  Row w[G][4];
  // row y enters the window slot q
  auto enter = [&](int y, int q) {
    const bool yin = y >= 0 && y < NY;
    const long c = o + x + (long)NX * (yin ? y : 0);
    const bool upd = yin && y >= 1 + isub && y <= NY - 2 - isub && (flags[c - o] & MB_VUPD);
    CellK ck{0.0f, 0.0f};
    if (upd) ck = cell_factors(Dio[c], Dn[c], dtc);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const float l = lam_in[tb[g] + c];
      w[g][q].f = f[tb[g] + c];
      w[g][q].acc = upd ? __builtin_fmaf(l, ck.am1, l) : l;
      w[g][q].mu = upd ? -ck.b * l : 0.0f;
    }
  };
  enter(y0 - 3, 0);
  enter(y0 - 2, 1);
  enter(y0 - 1, 2);
  enter(y0, 3);
  for (int Y = y0 - 2; Y <= y1; Y++) {
    FaceC<false> fc{};
    if (Y >= isub && Y <= NY - 2 - isub) {
      const long y2 = x + (long)NX * Y;
      fc = face_at<SCH>(s, Y, y2, o, flags, vv, hv, fk);
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      const float Fbar = w[g][1].mu - w[g][2].mu;
      float gr[4];
      face_flux_grad32<SCH, DIFF>(fc, w[g][0].f, w[g][1].f, w[g][2].f, w[g][3].f, gr);
#pragma unroll
      for (int q = 0; q < 4; q++) w[g][q].acc = __builtin_fmaf(Fbar, gr[q], w[g][q].acc);
    }
    const int yd = Y - 1;  // complete
    if (yd >= y0 && yd < y1) {
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) lam_out[tb[g] + o + x + (long)NX * yd] = w[g][0].acc;
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      w[g][0] = w[g][1];
      w[g][1] = w[g][2];
      w[g][2] = w[g][3];
    }
    enter(Y + 3, 3);
  }
}

// ---- transpose of the vertical sub-step ----------------------------------------
// one thread per column and group of kVGroup tracers (blockIdx.z).  Marching
// over the faces K' = 0..kmax-1, the window holds levels K'-1..K'+2 (the
// flux at K'+1/2 reads them): each level enters with its diagonal part (the
// adjoint through its own value) and mu = -b*lam (the adjoint of the flux
// through its upper face, into its own update), collects the face terms, and
// is stored when it leaves.  A face is the wk of level K' and the wprev of
// level K'+1 (K' >= 1, or K' = 0 with kshift).  The face constants and the
// cell factors serve the whole group.
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
  constexpr int G = kVGroup;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int t0 = blockIdx.z * G;
  if (i >= d.NX || j >= d.NY || t0 >= ntr) return;
  const int kmax = d.kmax;
  const long NXY = d.NXY, c2 = i + (long)d.NX * j;
  long tb[G];
#pragma unroll
  for (int g = 0; g < G; g++) tb[g] = (long)min(t0 + g, ntr - 1) * ts;
  const bool wcol = (flags[c2] & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2;
  if (!wcol) {
    for (int K = 0; K <= kmax; K++)
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) lam_out[tb[g] + c2 + NXY * K] = lam_in[tb[g] + c2 + NXY * K];
    return;
  }
  struct Lev {
    float f, acc, mu;
  };
  Lev w[G][4];
  auto enter = [&](int L, int q) {
    const bool in = L <= kmax;
    const long cL = c2 + NXY * (in ? L : kmax);
    const bool kdo = in && L >= 1 && L <= kmax - kshift;
    CellK ck{0.0f, 0.0f};
    if (kdo) ck = cell_factors(D2[cL], D3[cL], dt);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const float l = lam_in[tb[g] + cL];
      w[g][q].f = f[tb[g] + cL];
      w[g][q].acc = kdo ? __builtin_fmaf(l, ck.am1, l) : l;
      w[g][q].mu = kdo ? -ck.b * l : 0.0f;
    }
  };
#pragma unroll
  for (int g = 0; g < G; g++) w[g][0] = Lev{0.0f, 0.0f, 0.0f};
  enter(0, 1);
  enter(1, 2);
  enter(2, 3);
  double d0 = D2[c2], d1 = D2[c2 + NXY];  // D2 at K', K'+1
  for (int Kf = 0; Kf <= kmax - 1; Kf++) {
    // level 0 is never updated (mu = 0), so at K' = 0 Fbar is -mu(1) with
    // kshift and 0 without
    const bool used = Kf >= 1 || kshift;
    WFace<false> wf;
    wface_setup<SCH>(wf, ww[c2 + NXY * Kf], dt, d0, d1);
    const bool up = wf.w > 0.0f;
    const bool lim = up ? (SCH != UPSTREAM && Kf > 1 - kshift)
                        : (SCH != UPSTREAM && Kf < kmax - 1);
    const float gl = lim ? wf.g : 0.0f;
#pragma unroll
    for (int g = 0; g < G; g++) {
      const float Fbar = used ? w[g][1].mu - w[g][2].mu : 0.0f;
      const float fkm1 = Kf >= 1 ? w[g][0].f : 0.0f, fk = w[g][1].f, fk1 = w[g][2].f;
      const float fk2 = Kf + 2 <= kmax ? w[g][3].f : fk;
      const float fu = up ? fk : fk1, fuu = up ? fkm1 : fk2, fd = up ? fk1 : fk;
      const RecGrad32 r = rec_fast_grad<SCH>(fuu, fu, fd, gl, wf.A, wf.al, wf.E);
      const float a = Fbar * wf.w;
      // (a slot past either end of the column is never stored)
      const float a0 = up ? a * r.duu : 0.0f, a1 = up ? a * r.du : a * r.dd;
      const float a2 = up ? a * r.dd : a * r.du, a3 = up ? 0.0f : a * r.duu;
      w[g][0].acc += a0;
      w[g][1].acc += a1;
      w[g][2].acc += a2;
      w[g][3].acc += a3;
    }
    if (Kf >= 1) {
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) lam_out[tb[g] + c2 + NXY * (Kf - 1)] = w[g][0].acc;
    }
#pragma unroll
    for (int g = 0; g < G; g++) {
      w[g][0] = w[g][1];
      w[g][1] = w[g][2];
      w[g][2] = w[g][3];
    }
    enter(Kf + 3, 3);
    d0 = d1;
    d1 = Kf + 2 <= kmax ? D2[c2 + NXY * (Kf + 2)] : 0.0;
  }
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (t0 + g < ntr) {
      lam_out[tb[g] + c2 + NXY * (kmax - 1)] = w[g][0].acc;
      lam_out[tb[g] + c2 + NXY * kmax] = w[g][1].acc;
    }
  }
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
    if (dir == 0) {
      hipLaunchKernelGGL((k_sub_u_f32_adj<SCH, DIFF>),
                         dim3((d.NX + kUOut - 1) / kUOut, (d.NY + 3) / 4, d.kmax + 1),
                         dim3(64, 4), 0, st, d, ntr, ts, a, b, out, Dio, Dn, vel, Dv, flags, dt,
                         lf, lc, ard1, AH);
    } else {
      const int ng = (ntr + kVGroup - 1) / kVGroup;
      hipLaunchKernelGGL((k_sub_v_f32_adj<SCH, DIFF>),
                         dim3((d.NX + 63) / 64, (d.NY + kVRows - 1) / kVRows, (d.kmax + 1) * ng),
                         dim3(64), 0, st, d, ntr, ts, a, b, out, Dio, Dn, vel, Dv, flags, dt, lc,
                         lf, ard1, AH, ng);
    }
  } else {
    (void)b;
    if (dir == 0) {
      hipLaunchKernelGGL((k_sub_u_f32<SCH, DIFF>),
                         dim3((d.NX + kUFwdOut - 1) / kUFwdOut, (d.NY + 3) / 4, d.kmax + 1),
                         dim3(64, 4), 0, st, d, ntr, ts, a, out, Dio, Dn, vel, Dv, flags, dt, lf,
                         lc, ard1, AH);
    } else {
      const int ng = (ntr + kVGroup - 1) / kVGroup;
      hipLaunchKernelGGL((k_sub_v_f32<SCH, DIFF>),
                         dim3((d.NX + 63) / 64, (d.NY + kVRows - 1) / kVRows, (d.kmax + 1) * ng),
                         dim3(64), 0, st, d, ntr, ts, a, out, Dio, Dn, vel, Dv, flags, dt, lc, lf,
                         ard1, AH, ng);
    }
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
    hipLaunchKernelGGL((k_sub_w_f32_adj<SCH>), grid3(d, (ntr + kVGroup - 1) / kVGroup),
                       dim3(64, 4), 0, st, d, ntr, ts, a, b,
                       out, D2, D3, ww, flags, dt, kshift);
  } else {
    (void)b;
    hipLaunchKernelGGL((k_sub_w_f32<SCH>), grid3(d, (ntr + kVGroup - 1) / kVGroup), dim3(64, 4),
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
