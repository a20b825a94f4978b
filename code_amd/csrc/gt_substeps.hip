// gt_substeps.hip -- the HALFSPLIT step one sub-step at a time, in both
// arithmetic flavours of the strip kernels, and the transpose of each
// sub-step: the reverse mode of the advection.
//
// There is no counterpart in getm-model/code; the forward operator is the
// HALFSPLIT composition of advection_3d.F90:439-547 (u/2, v/2, w, v/2, u/2)
// with the sub-steps of adv_split_u.F90, adv_split_v.F90 and
// adv_split_w.F90.
//
// Forward: one kernel per sub-step, built from the strip kernels' device
// functions (gt_strip_impl.h) with the same operands, so every intermediate
// state is the one the strip kernels hold at that point and the final state
// equals theirs bit for bit.  TF = double: the reference's expressions
// (EXACT); TF = float: the batched FP32 arithmetic.  The layer heights D1..D5
// depend on the velocities only and are formed once (k_heights), in FP64
// with the sub-steps' expressions.
//
// Reverse: with velocities and layer heights held fixed each sub-step is a
// piecewise linear map of the tracer,
//   f_new(i) = a(i) f(i) - s(i) (F(i) - F(i-1)),
//   F(X)     = vel(X) rec(fuu, fu, fd) - kA(X) (f(X+1) - f(X)),
// with a = Dio/Dn and s = dt*splitfac*dlen*ard1/Dn (FP32: 1+am1 and b, the
// factors of cell_update_fast).  The transpose is formed face by face, every
// face once: along i a lane owns a face and the cells collect their faces'
// terms by cross-lane moves; along j and k a thread marches rows or levels
// with a four-deep window in registers and stores each output when no
// further face reaches it (no atomics).  The limiter's derivative is that of
// the branch the forward took (one-sided at ties and kinks).
#include "gt_strip_impl.h"
#include "gt_substeps.h"

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

// ---- shared pieces -------------------------------------------------------------
template <typename TF>
constexpr bool kExact = sizeof(TF) == 8;
// tracers per thread of the row and column marches
template <typename TF>
constexpr int kGroup = sizeof(TF) == 8 ? 1 : 6;
// output rows per wavefront of the v marches (FP64 carries one tracer, so
// shorter marches for more wavefronts)
template <typename TF>
constexpr int kVRows = sizeof(TF) == 8 ? 8 : 16;
constexpr long kWaveTarget = 8192;  // wavefronts a column march aims for (8 per SIMD)
constexpr int kUOutF = 60;  // outputs per wavefront of the u forward
constexpr int kUOutA = 58;  // ... and of the u transpose

// geometry of one sweep: dir 0 = u (faces along i, stride 1), 1 = v (NX)
struct Sweep {
  long st;
  int n;
  uint8_t mflux, mupd;
};
__device__ __forceinline__ Sweep sweep(const Dims& d, int dir) {
  return dir == 0 ? Sweep{1, d.NX, MB_UFLUX, MB_UUPD} : Sweep{(long)d.NX, d.NY, MB_VFLUX, MB_VUPD};
}

// the constants of face X (between cells X and X+1 of the sweep) as the strip
// kernels form them; X in 0..n-2, x2 its column index, o its level offset
template <int SCH, bool EXACT>
__device__ __forceinline__ FaceC<EXACT> face_at(const Sweep& s, int X, long x2, long o,
                                                const uint8_t* flags, const double* vel,
                                                const double* Dv, const FaceK& k) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const bool active = X >= isub && X <= s.n - 2 - isub && (flags[x2] & s.mflux);
  const bool lim_l = SCH != UPSTREAM && X >= 1 && (flags[x2 - s.st] & s.mflux);
  const bool lim_r = SCH != UPSTREAM && (flags[x2 + s.st] & s.mflux);
  FaceC<EXACT> fc;
  face_setup<SCH>(fc, active, lim_l, lim_r, vel[o + x2], Dv[o + x2], k);
  return fc;
}

// a face that carries no flux (outside the sweep's face range)
template <bool EXACT>
__device__ __forceinline__ FaceC<EXACT> no_face() {
  FaceC<EXACT> fc{};
  if constexpr (EXACT) fc.code = F_NONE;
  return fc;
}

// the loads of row Y of a v march: face Y's transport, layer height and
// flux-mask bytes (rows Y-1..Y+1) and cell Y's heights before and after,
// read one row ahead of their use (rows outside 0..NY-1 read a clamped row;
// the face and the update are switched off there by Y itself)
struct VRaw {
  double vel, Dv, Dio, Dn;
  uint8_t fm, f0, fp;
};
__device__ __forceinline__ VRaw vraw(const Dims& d, int x, long o, int Y,
                                     const uint8_t* flags, const double* vel,
                                     const double* Dv, const double* Dio, const double* Dn) {
  const int NX = d.NX, Yc = min(max(Y, 0), d.NY - 1);
  const long x2 = x + (long)NX * Yc;
  VRaw r;
  r.vel = vel[o + x2];
  r.Dv = Dv[o + x2];
  r.Dio = Dio[o + x2];
  r.Dn = Dn[o + x2];
  r.f0 = flags[x2];
  r.fm = flags[x + (long)NX * max(Yc - 1, 0)];
  r.fp = flags[x + (long)NX * min(Yc + 1, d.NY - 1)];
  return r;
}
// face Y of the v sweep from its loads, as face_at forms it (0 <= Y <= NY-2)
template <int SCH, bool EXACT>
__device__ __forceinline__ FaceC<EXACT> vface(const Dims& d, int Y, const VRaw& r,
                                              const FaceK& k) {
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const bool active = Y >= isub && Y <= d.NY - 2 - isub && (r.f0 & MB_VFLUX);
  const bool lim_l = SCH != UPSTREAM && Y >= 1 && (r.fm & MB_VFLUX);
  const bool lim_r = SCH != UPSTREAM && (r.fp & MB_VFLUX);
  FaceC<EXACT> fc;
  face_setup<SCH>(fc, active, lim_l, lim_r, r.vel, r.Dv, k);
  return fc;
}

// the update of one cell: f_new = (Dio f - dt advn)/Dn (EXACT, the
// reference's expression) or f + f am1 - (Fc - Fm) b (FP32), and its
// linearisation: diag (d f_new / d f) and mu (d f_new / d Fc, = -d/d Fm)
template <bool EXACT>
struct CellOp;
template <>
struct CellOp<false> {
  float am1 = 0.0f, b = 0.0f;
  // dtc = dt*splitfac*dlen*ard1 (dt for w)
  __device__ static CellOp make(double Dio, double Dn, double dtc) {
    const CellK k = cell_factors(Dio, Dn, dtc);
    CellOp c;
    c.am1 = k.am1;
    c.b = k.b;
    return c;
  }
  __device__ float h_step(float fo, float Fc, float Fm, double, double, double) const {
    return cell_update_fast(CellK{am1, b}, fo, Fc, Fm);
  }
  __device__ float w_step(float fo, float wk, float wprev, double) const {
    return cell_update_fast(CellK{am1, b}, fo, wk, wprev);
  }
  __device__ float diag(float l) const { return __builtin_fmaf(l, am1, l); }
  __device__ float mu(float l, double, double) const { return -b * l; }
};
template <>
struct CellOp<true> {
  double Dio = 1.0, Dn = 1.0;
  __device__ static CellOp make(double Dio, double Dn, double) {
    CellOp c;
    c.Dio = Dio;
    c.Dn = Dn;
    return c;
  }
  // adv_split_u.F90:267-286 in the reference's order (dlen: the face length)
  __device__ double h_step(double fo, double Fc, double Fm, double dt, double dlen,
                           double ard1) const {
    double advn;
    return cell_update_exact(true, Dio, Dn, fo, Fc, Fm, dlen, ard1, dt, &advn);
  }
  // adv_split_w.F90:287-294
  __device__ double w_step(double fo, double wk, double wprev, double dt) const {
    const double aw = 1.0 * (wk - wprev);
    return (Dio * fo - dt * aw) / Dn;
  }
  __device__ double diag(double l) const { return l * Dio / Dn; }
  // fac = splitfac*dlen*ard1 (splitfac for w)
  __device__ double mu(double l, double dt, double fac) const { return l * (-dt / Dn) * fac; }
};

// ---- derivatives of the reconstruction --------------------------------------------
template <typename T>
struct RecG {
  T duu, du, dd;
};

// the reference arithmetic (reconstruct_bf): rec = fu + (1-cfl)/2 phi(r)
// deltaf, r = deltafu/deltaf, the branch of phi that the forward's min / max
// selected; E = 2/(1-cfl) as the face holds it
template <int SCH>
__device__ __forceinline__ RecG<double> rec_grad_exact(double cfl, double E, double fuu, double fu,
                                                       double fd) {
  const double g = 0.5 * (1.0 - cfl);
  const double deltaf = fd - fu, deltafu = fu - fuu;
  if constexpr (SCH == UPSTREAM) {
    return {0.0, 1.0, 0.0};
  } else if constexpr (SCH == P2) {  // linear on both branches
    const double x = (1.0 / 6.0) * (1.0 - 2.0 * cfl);
    return {-g * (0.5 - x), 1.0 - g * (0.5 + x) + g * (0.5 - x), g * (0.5 + x)};
  } else {
    if (!(deltaf * deltafu > 0.0)) return {0.0, 1.0, 0.0};
    const double r = deltafu / deltaf;
    double phi, dphi;
    if constexpr (SCH == P2_PDM) {
      const double x = (1.0 / 6.0) * (1.0 - 2.0 * cfl);
      const double a = 2.0 * r / (cfl + 1.e-10), b = (0.5 + x) + (0.5 - x) * r;
      const double m = a < b ? a : b;
      const double dm = a < b ? 2.0 / (cfl + 1.e-10) : (0.5 - x);
      if (E < m) { phi = E; dphi = 0.0; } else { phi = m; dphi = dm; }
    } else if constexpr (SCH == SUPERBEE) {
      const double a = 2.0 * r < 1.0 ? 2.0 * r : 1.0, da = 2.0 * r < 1.0 ? 2.0 : 0.0;
      const double b = r < 2.0 ? r : 2.0, db = r < 2.0 ? 1.0 : 0.0;
      if (a > b) { phi = a; dphi = da; } else { phi = b; dphi = db; }
    } else {  // MUSCL
      const double a = 2.0 * r, b = 0.5 * (1.0 + r);
      const double m = a < b ? a : b, dm = a < b ? 2.0 : 0.5;
      if (2.0 < m) { phi = 2.0; dphi = 0.0; } else { phi = m; dphi = dm; }
    }
    // d/dfuu = -g dphi, d/dfu = 1 + g (dphi (1 + r) - phi), d/dfd = g (phi - dphi r)
    return {-g * dphi, 1.0 + g * (dphi * (1.0 + r) - phi), g * (phi - dphi * r)};
  }
}

// the FP32 arithmetic (rec_fast): rec = fu + g t, t the branch term the
// forward's min / max selected, each term linear in du = fu-fuu, df = fd-fu;
// the identity where the limiter does not apply (g = 0) or deltaf*deltafu <= 0
template <int SCH>
__device__ __forceinline__ RecG<float> rec_grad_fast(float fuu, float fu, float fd, float g,
                                                     float A, float al, float E) {
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
      // whether rec_fast's fminf (lo) / fmaxf (hi) of (a, b) returns b
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
    return {-g * tu, 1.0f + g * (tu - tf), g * tf};
  }
}

// partials of face_flux with respect to the cells X-1..X+2
template <typename TF, int SCH, bool DIFF>
__device__ __forceinline__ void face_flux_grad(const FaceC<kExact<TF>>& fc, TF fm1, TF f0, TF fp1,
                                               TF fp2, double AH, double dlen, TF out[4]) {
  if constexpr (kExact<TF>) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (fc.code == F_NONE) return;
    const bool left = fc.code <= F_LEFT_LIM;
    const bool lim = fc.code == F_LEFT_LIM || fc.code == F_RIGHT_LIM;
    const double fuu = left ? fm1 : fp2, fu0 = left ? f0 : fp1, fd = left ? fp1 : f0;
    const RecG<double> r =
        lim ? rec_grad_exact<SCH>(fc.cfl, fc.E, fuu, fu0, fd) : RecG<double>{0.0, 1.0, 0.0};
    if (left) {
      out[0] = fc.vel * r.duu; out[1] = fc.vel * r.du; out[2] = fc.vel * r.dd;
    } else {
      out[3] = fc.vel * r.duu; out[2] = fc.vel * r.du; out[1] = fc.vel * r.dd;
    }
    if constexpr (DIFF) {
      const double k = AH * fc.Dv / dlen;
      out[1] += k;
      out[2] -= k;
    }
  } else {
    (void)AH; (void)dlen;
    const float fuu = fc.right ? fp2 : fm1, fu0 = fc.right ? fp1 : f0, fd = fc.right ? f0 : fp1;
    const RecG<float> r = rec_grad_fast<SCH>(fuu, fu0, fd, fc.g, fc.A, fc.al, fc.E);
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
}

// the w flux at face K (K+1/2) from the levels K-1..K+2, and its partials;
// lim_up / lim_dn: whether the limiter applies going up / down
template <typename TF, int SCH>
__device__ __forceinline__ TF w_flux(const WFace<kExact<TF>>& wf, int K, int kshift, int kmax,
                                     bool lim_up, bool lim_dn, TF fkm1, TF fk, TF fk1, TF fk2) {
  if constexpr (kExact<TF>) {
    (void)lim_up; (void)lim_dn;
    return wflux<TF, SCH, true>(wf, K, kshift, kmax, fkm1, fk, fk1, fk2);
  } else {
    (void)K; (void)kshift; (void)kmax;
    return wflux_fast<SCH>(wf, lim_up, lim_dn, fkm1, fk, fk1, fk2);
  }
}

template <typename TF, int SCH>
__device__ __forceinline__ void w_flux_grad(const WFace<kExact<TF>>& wf, bool lim_up, bool lim_dn,
                                            TF fkm1, TF fk, TF fk1, TF fk2, TF out[4]) {
  const bool up = wf.w > 0;
  const TF fu = up ? fk : fk1, fuu = up ? fkm1 : fk2, fd = up ? fk1 : fk;
  const bool lim = up ? lim_up : lim_dn;
  RecG<TF> r;
  if constexpr (kExact<TF>) {
    r = lim ? rec_grad_exact<SCH>(wf.cfl, wf.E, fuu, fu, fd) : RecG<double>{0.0, 1.0, 0.0};
  } else {
    r = rec_grad_fast<SCH>(fuu, fu, fd, lim ? wf.g : 0.0f, wf.A, wf.al, wf.E);
  }
  const TF w = TF(wf.w);
  if (up) {
    out[0] = w * r.duu; out[1] = w * r.du; out[2] = w * r.dd; out[3] = TF(0);
  } else {
    out[0] = TF(0); out[1] = w * r.dd; out[2] = w * r.du; out[3] = w * r.duu;
  }
}

// ---- the u half sub-step ---------------------------------------------------------
// A wavefront holds 64 consecutive cells of a row; lane L forms the flux of
// face x from the cells x-1..x+2 (cross-lane moves) and cell x takes the flux
// of face x-1 from lane L-1, as the strip kernels do: lanes 2..61 are
// complete.  Level 0 and cells the sub-step leaves alone are copied.
template <typename TF, int SCH, bool DIFF>
__global__ __launch_bounds__(256) void k_sub_u(Dims d, int ntr, long ts,
                                               const TF* __restrict__ fin, TF* __restrict__ fout,
                                               const double* __restrict__ Dio,
                                               const double* __restrict__ Dn,
                                               const double* __restrict__ uu,
                                               const double* __restrict__ hu,
                                               const uint8_t* __restrict__ flags, double dt,
                                               double dx, double dy, double ard1, double AH) {
  constexpr bool EX = kExact<TF>;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const int lane = threadIdx.x;  // blockDim (64, 4): one row per wavefront
  const int x = blockIdx.x * kUOutF - 2 + lane;
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
  FaceC<EX> fc = no_face<EX>();
  if (xin && x <= d.NX - 2)
    fc = face_at<SCH, EX>(sweep(d, 0), x, c2, o, flags, uu, hu, face_k(0.5 * dt, dx, AH));
  const bool upd = xin && x >= 1 + isub && x <= d.NX - 2 - isub && (flags[c2] & MB_UUPD);
  CellOp<EX> cu;
  if (upd) cu = CellOp<EX>::make(Dio[c], Dn[c], dt * (0.5 * dy * ard1));
  for (int t = 0; t < ntr; t++) {
    const TF fx = fin[t * ts + c];
    const TF fm1 = from_lower_lane(fx), fp1 = from_upper_lane(fx);
    const TF fp2 = from_upper_lane(fp1);
    const TF Fc = face_flux<TF, SCH, DIFF>(fc, fm1, fx, fp1, fp2, AH, dx);
    const TF Fm = from_lower_lane(Fc);
    const TF u = cu.h_step(fx, Fc, Fm, dt, dy, ard1);
    if (out) fout[t * ts + c] = upd ? u : fx;
  }
}

// ---- the v half sub-step: a row march -------------------------------------------
// A wavefront holds 64 columns and marches the faces Y = y0-1 .. y1-1 for a
// group of tracers, with the rows Y-1..Y+2 and the flux of face Y-1 in
// registers; row Y is updated after face Y.
template <typename TF, int SCH, bool DIFF>
__global__ __launch_bounds__(64) void k_sub_v(Dims d, int ntr, long ts, const TF* __restrict__ fin,
                                              TF* __restrict__ fout,
                                              const double* __restrict__ Dio,
                                              const double* __restrict__ Dn,
                                              const double* __restrict__ vv,
                                              const double* __restrict__ hv,
                                              const uint8_t* __restrict__ flags, double dt,
                                              double dx, double dy, double ard1, double AH,
                                              int ngroups) {
  constexpr bool EX = kExact<TF>;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  constexpr int G = kGroup<TF>;
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y0 = blockIdx.y * kVRows<TF>;
  const int k = blockIdx.z / ngroups, t0 = (blockIdx.z % ngroups) * G;
  if (x >= d.NX) return;  // no cross-lane moves here
  const int NX = d.NX, NY = d.NY;
  const long o = d.NXY * (long)k;
  const int y1 = min(y0 + kVRows<TF>, NY);
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
  const FaceK fk = face_k(0.5 * dt, dy, AH);
  const double dtc = dt * (0.5 * dx * ard1);
  auto row = [&](int y) -> long { return o + x + (long)NX * (y >= 0 && y < NY ? y : 0); };
  TF w[G][4], Fp[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    Fp[g] = TF(0);
#pragma unroll
    for (int q = 0; q < 4; q++) w[g][q] = fin[tb[g] + row(y0 - 2 + q)];
  }
  VRaw rn = vraw(d, x, o, y0 - 1, flags, vv, hv, Dio, Dn);
  for (int Y = y0 - 1; Y < y1; Y++) {
    const VRaw r = rn;
    rn = vraw(d, x, o, Y + 1, flags, vv, hv, Dio, Dn);
    FaceC<EX> fc = no_face<EX>();
    if (Y >= 0 && Y <= NY - 2) fc = vface<SCH, EX>(d, Y, r, fk);
    const bool st = Y >= y0;
    const long cY = o + x + (long)NX * (Y >= 0 ? Y : 0);
    const bool upd = st && Y >= 1 + isub && Y <= NY - 2 - isub && (r.f0 & MB_VUPD);
    CellOp<EX> cu;
    if (upd) cu = CellOp<EX>::make(r.Dio, r.Dn, dtc);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const TF F = face_flux<TF, SCH, DIFF>(fc, w[g][0], w[g][1], w[g][2], w[g][3], AH, dy);
      if (st && t0 + g < ntr) {
        const TF u = cu.h_step(w[g][1], F, Fp[g], dt, dx, ard1);
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

// ---- the vertical sub-step (adv_ver_iterations = 1) ------------------------------
// One thread per column, group of tracers and chunk of levels, marching up with
// the levels K-1..K+2 of each tracer, K..K+3 of D2, D3 and ww and the flux
// at K-1/2 in registers (kmax >= 2).  At level K, as the strip kernels set them
// (adv_split_w.F90:249-300): the flux at K+1/2 exists below kmax, the one at
// K-1/2 above level 1 (or at level 1 with kshift); level K is updated up to
// kmax-kshift; the limiter applies above level 1-kshift going up and below
// kmax-1 going down.
template <typename TF, int SCH>
__global__ __launch_bounds__(256) void k_sub_w(Dims d, int ntr, long ts, const TF* __restrict__ fin,
                                               TF* __restrict__ fout,
                                               const double* __restrict__ D2,
                                               const double* __restrict__ D3,
                                               const double* __restrict__ ww,
                                               const uint8_t* __restrict__ flags, double dt,
                                               int kshift, int ngroups, int nchunks, int kc) {
  constexpr bool EX = kExact<TF>;
  constexpr int G = kGroup<TF>;
  // blockIdx.x = column block * nz + z, z = chunk * ngroups + group: the
  // groups and chunks of one column block run together, so the FP64 fields
  // they all read meet in the cache
  const int nz = ngroups * nchunks, bz = (int)(blockIdx.x % nz);
  const int i = (int)(blockIdx.x / nz) * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int t0 = (bz % ngroups) * G;
  const int kmax = d.kmax;
  // this thread's levels K0..K1 (chunk bz / ngroups of kc levels; the first
  // chunk also carries level 0)
  const int K0 = 1 + (bz / ngroups) * kc, K1 = min(K0 + kc - 1, kmax);
  if (i >= d.NX || j >= d.NY || t0 >= ntr || K0 > kmax) return;
  const long NXY = d.NXY, c2 = i + (long)d.NX * j;
  long tb[G];
#pragma unroll
  for (int g = 0; g < G; g++) tb[g] = (long)min(t0 + g, ntr - 1) * ts;
  const bool wcol = (flags[c2] & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2;
  if (!wcol) {
    for (int K = K0 == 1 ? 0 : K0; K <= K1; K++)
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) fout[tb[g] + c2 + NXY * K] = fin[tb[g] + c2 + NXY * K];
    return;
  }
  TF n[G][5];  // n[1..4]: levels K-1..K+2 (entries past kmax unused)
  TF fKm2[G];  // level K0-2 (zero below level 0), for the flux at K0-1/2
#pragma unroll
  for (int g = 0; g < G; g++) {
    const TF* f = fin + tb[g];
    fKm2[g] = K0 >= 2 ? f[c2 + NXY * (K0 - 2)] : TF(0);
    n[g][1] = f[c2 + NXY * (K0 - 1)];
    n[g][2] = f[c2 + NXY * K0];
    n[g][3] = K0 + 1 <= kmax ? f[c2 + NXY * (K0 + 1)] : TF(0);
    n[g][4] = K0 + 2 <= kmax ? f[c2 + NXY * (K0 + 2)] : TF(0);
    if (K0 == 1 && t0 + g < ntr) fout[tb[g] + c2] = n[g][1];  // level 0
  }
  // D2, D3 and ww of levels K..K+3 in registers, each loaded three levels
  // ahead of its use (levels past kmax read level kmax, never used)
  auto lv = [&](int L) -> long { return c2 + NXY * (long)min(L, kmax); };
  double d2r[4], d3r[4], wr[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    d2r[q] = D2[lv(K0 + q)];
    d3r[q] = D3[lv(K0 + q)];
    wr[q] = ww[lv(K0 + q)];
  }
  // The flux at K-1/2 that level K reads is the flux at K+1/2 of level K-1,
  // so each face is evaluated once and carried up (pc).  The operands agree
  // except the one past kmax that the strip kernels substitute, which is the
  // upwind-upwind level of a downward face with the limiter off: not read.
  TF pc[G];
  {
    const int F = K0 - 1;  // the face below level K0
    WFace<EX> wf0;
    wface_setup<SCH>(wf0, ww[c2 + NXY * F], dt, D2[c2 + NXY * F], d2r[0]);
    const bool lim_up = SCH != UPSTREAM && F > 1 - kshift;
    const bool lim_dn = SCH != UPSTREAM && F < kmax - 1;
#pragma unroll
    for (int g = 0; g < G; g++)
      pc[g] = w_flux<TF, SCH>(wf0, F, kshift, kmax, lim_up, lim_dn, fKm2[g], n[g][1], n[g][2],
                              K0 + 1 <= kmax ? n[g][3] : n[g][2]);
  }
  for (int K = K0; K <= K1; K++) {
    const long cK = c2 + NXY * K;
    const bool wk_on = K <= kmax - 1, wp_on = K > 1 || kshift, kdo = K <= kmax - kshift;
    const bool limk_up = SCH != UPSTREAM && K > 1 - kshift;
    const bool limk_dn = SCH != UPSTREAM && K < kmax - 1;
    const double dk = d2r[0], dp = d2r[1];  // D2 at K, K+1
    WFace<EX> wfk;
    wface_setup<SCH>(wfk, wr[0], dt, dk, K + 1 <= kmax ? dp : dk);
    const CellOp<EX> cw = CellOp<EX>::make(dk, d3r[0], dt);
    const bool nxt = K + 3 <= kmax;
    const long cn = c2 + NXY * (nxt ? K + 3 : kmax);
#pragma unroll
    for (int g = 0; g < G; g++) {
      // the operands of the strip kernels: levels past kmax are replaced by
      // level K
      const TF fKm1 = n[g][1], fK = n[g][2];
      const TF fKp1 = K + 1 <= kmax ? n[g][3] : fK, fKp2 = K + 2 <= kmax ? n[g][4] : fK;
      const TF q = w_flux<TF, SCH>(wfk, K, kshift, kmax, limk_up, limk_dn, fKm1, fK, fKp1, fKp2);
      const TF wprev = wp_on ? pc[g] : TF(0), wk = wk_on ? q : TF(0);
      const TF r = kdo ? cw.w_step(fK, wk, wprev, dt) : fK;
      if (t0 + g < ntr) fout[tb[g] + cK] = r;
      pc[g] = q;
      n[g][1] = n[g][2];
      n[g][2] = n[g][3];
      n[g][3] = n[g][4];
      n[g][4] = nxt ? fin[tb[g] + cn] : TF(0);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      d2r[q] = d2r[q + 1];
      d3r[q] = d3r[q + 1];
      wr[q] = wr[q + 1];
    }
    d2r[3] = D2[lv(K + 4)];
    d3r[3] = D3[lv(K + 4)];
    wr[3] = ww[lv(K + 4)];
  }
}

// ---- transpose of the u sub-step: one face per lane, cross-lane sums --------
// Lane L owns cell x and face x (between x and x+1).  Face x reads the cells
// x-1..x+2 and its flux enters the updates of cells x and x+1; cell x
// collects the faces x-2..x+1 (lanes L-2..L+1).  Lanes 3..60 are complete.
// f: the state before the sub-step; lam_in: the adjoint of its output;
// lam_out: of its input.
template <typename TF, int SCH, bool DIFF>
__global__ __launch_bounds__(256) void k_sub_u_adj(
    Dims d, int ntr, long ts, const TF* __restrict__ f, const TF* __restrict__ lam_in,
    TF* __restrict__ lam_out, const double* __restrict__ Dio, const double* __restrict__ Dn,
    const double* __restrict__ uu, const double* __restrict__ hu,
    const uint8_t* __restrict__ flags, double dt, double dx, double dy, double ard1, double AH) {
  constexpr bool EX = kExact<TF>;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const int lane = threadIdx.x;
  const int x = blockIdx.x * kUOutA - 3 + lane;
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
  FaceC<EX> fc = no_face<EX>();
  if (xin && x <= d.NX - 2)
    fc = face_at<SCH, EX>(sweep(d, 0), x, c2, o, flags, uu, hu, face_k(0.5 * dt, dx, AH));
  const bool upd = xin && x >= 1 + isub && x <= d.NX - 2 - isub && (flags[c2] & MB_UUPD);
  const double fac = 0.5 * dy * ard1;
  CellOp<EX> cu;
  if (upd) cu = CellOp<EX>::make(Dio[c], Dn[c], dt * fac);
  for (int t = 0; t < ntr; t++) {
    const TF fx = f[t * ts + c], lx = lam_in[t * ts + c];
    const TF fm1 = from_lower_lane(fx), fp1 = from_upper_lane(fx);
    const TF fp2 = from_upper_lane(fp1);
    const TF mu = upd ? cu.mu(lx, dt, fac) : TF(0);  // d(out x)/d F(x) times lam
    const TF Fbar = mu - from_upper_lane(mu);
    TF g[4];
    face_flux_grad<TF, SCH, DIFF>(fc, fm1, fx, fp1, fp2, AH, dx, g);
    TF acc = upd ? cu.diag(lx) : lx;
    acc += Fbar * g[1];                                   // face x
    acc += from_upper_lane(Fbar * g[0]);                  // face x+1
    acc += from_lower_lane(Fbar * g[2]);                  // face x-1
    acc += from_lower_lane(from_lower_lane(Fbar * g[3])); // face x-2
    if (out) lam_out[t * ts + c] = acc;
  }
}

// ---- transpose of the v sub-step: a row march -----------------------------------
// The faces Y = y0-2 .. y1 (those of the output rows); before face Y the
// window holds rows Y-1..Y+2: each row enters with its diagonal part and mu,
// collects the terms of the faces whose stencil holds it, and is stored when
// it leaves (row Y-1 after face Y).
template <typename TF, int SCH, bool DIFF>
__global__ __launch_bounds__(64) void k_sub_v_adj(
    Dims d, int ntr, long ts, const TF* __restrict__ f, const TF* __restrict__ lam_in,
    TF* __restrict__ lam_out, const double* __restrict__ Dio, const double* __restrict__ Dn,
    const double* __restrict__ vv, const double* __restrict__ hv,
    const uint8_t* __restrict__ flags, double dt, double dx, double dy, double ard1, double AH,
    int ngroups) {
  constexpr bool EX = kExact<TF>;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  constexpr int G = kGroup<TF>;
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y0 = blockIdx.y * kVRows<TF>;
  const int k = blockIdx.z / ngroups, t0 = (blockIdx.z % ngroups) * G;
  if (x >= d.NX) return;  // no cross-lane moves here
  const int NX = d.NX, NY = d.NY;
  const long o = d.NXY * (long)k;
  const int y1 = min(y0 + kVRows<TF>, NY);
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
  const FaceK fk = face_k(0.5 * dt, dy, AH);
  const double fac = 0.5 * dx * ard1;
  struct Row {
    TF f, acc, mu;
  };
  Row w[G][4];
  // the loads of a row entering the window, read one face ahead
  struct ERaw {
    TF l[G], f[G];
    double Dio, Dn;
    uint8_t fl;
  };
  auto fetch = [&](int y) -> ERaw {
    const bool yin = y >= 0 && y < NY;
    const long cy = o + x + (long)NX * (yin ? y : 0);
    ERaw r;
    r.Dio = Dio[cy];
    r.Dn = Dn[cy];
    r.fl = flags[cy - o];
#pragma unroll
    for (int g = 0; g < G; g++) {
      r.l[g] = lam_in[tb[g] + cy];
      r.f[g] = f[tb[g] + cy];
    }
    return r;
  };
  // row y enters the window slot q
  auto enter = [&](int y, int q, const ERaw& r) {
    const bool yin = y >= 0 && y < NY;
    const bool upd = yin && y >= 1 + isub && y <= NY - 2 - isub && (r.fl & MB_VUPD);
    CellOp<EX> cu;
    if (upd) cu = CellOp<EX>::make(r.Dio, r.Dn, dt * fac);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const TF l = r.l[g];
      w[g][q].f = r.f[g];
      w[g][q].acc = upd ? cu.diag(l) : l;
      w[g][q].mu = upd ? cu.mu(l, dt, fac) : TF(0);
    }
  };
  enter(y0 - 3, 0, fetch(y0 - 3));
  enter(y0 - 2, 1, fetch(y0 - 2));
  enter(y0 - 1, 2, fetch(y0 - 1));
  enter(y0, 3, fetch(y0));
  ERaw en = fetch(y0 + 1);
  VRaw rn = vraw(d, x, o, y0 - 2, flags, vv, hv, Dio, Dn);
  for (int Y = y0 - 2; Y <= y1; Y++) {
    const VRaw r = rn;
    rn = vraw(d, x, o, Y + 1, flags, vv, hv, Dio, Dn);
    const ERaw e = en;
    en = fetch(Y + 4);
    FaceC<EX> fc = no_face<EX>();
    if (Y >= 0 && Y <= NY - 2) fc = vface<SCH, EX>(d, Y, r, fk);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const TF Fbar = w[g][1].mu - w[g][2].mu;
      TF gr[4];
      face_flux_grad<TF, SCH, DIFF>(fc, w[g][0].f, w[g][1].f, w[g][2].f, w[g][3].f, AH, dy, gr);
#pragma unroll
      for (int q = 0; q < 4; q++) w[g][q].acc += Fbar * gr[q];
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
    enter(Y + 3, 3, e);
  }
}

// ---- transpose of the vertical sub-step ----------------------------------------
// One thread per column and group of tracers.  Marching over the faces
// K' = 0..kmax-1, the window holds levels K'-1..K'+2 (the flux at K'+1/2
// reads them): each level enters with its diagonal part and mu, collects the
// face terms, and is stored when it leaves.  A face is the wk of level K' and
// the wprev of level K'+1 (K' >= 1, or K' = 0 with kshift).
template <typename TF, int SCH>
__global__ __launch_bounds__(256) void k_sub_w_adj(Dims d, int ntr, long ts,
                                                   const TF* __restrict__ f,
                                                   const TF* __restrict__ lam_in,
                                                   TF* __restrict__ lam_out,
                                                   const double* __restrict__ D2,
                                                   const double* __restrict__ D3,
                                                   const double* __restrict__ ww,
                                                   const uint8_t* __restrict__ flags, double dt,
                                                   int kshift, int ngroups, int nchunks, int kc) {
  constexpr bool EX = kExact<TF>;
  constexpr int G = kGroup<TF>;
  // blockIdx.x = column block * nz + z, z = chunk * ngroups + group: the
  // groups and chunks of one column block run together, so the FP64 fields
  // they all read meet in the cache
  const int nz = ngroups * nchunks, bz = (int)(blockIdx.x % nz);
  const int i = (int)(blockIdx.x / nz) * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int t0 = (bz % ngroups) * G;
  const int kmax = d.kmax;
  // this thread stores levels L0..L1 (chunk bz / ngroups of kc levels); it
  // marches the faces L0-2..L1+1 that reach them, in the order of the whole
  // column, so every stored sum is formed in the same order
  const int L0 = (bz / ngroups) * kc, L1 = min(L0 + kc - 1, kmax);
  if (i >= d.NX || j >= d.NY || t0 >= ntr || L0 > kmax) return;
  const int Fs = max(L0 - 2, 0), Fe = min(L1 + 1, kmax - 1);
  const long NXY = d.NXY, c2 = i + (long)d.NX * j;
  long tb[G];
#pragma unroll
  for (int g = 0; g < G; g++) tb[g] = (long)min(t0 + g, ntr - 1) * ts;
  const bool wcol = (flags[c2] & MB_WCOL) && i <= d.NX - 2 && j <= d.NY - 2;
  if (!wcol) {
    for (int K = L0; K <= L1; K++)
#pragma unroll
      for (int g = 0; g < G; g++)
        if (t0 + g < ntr) lam_out[tb[g] + c2 + NXY * K] = lam_in[tb[g] + c2 + NXY * K];
    return;
  }
  struct Lev {
    TF f, acc, mu;
  };
  Lev w[G][4];
  // a level's loads (read two faces before it enters the window; levels past
  // kmax read level kmax, whose slot is never stored)
  struct Raw {
    TF l[G], f[G];
    double d2, d3;
  };
  auto lv = [&](int L) -> long { return c2 + NXY * (long)min(L, kmax); };
  auto fetch = [&](int L, Raw& r) {
    const long cL = lv(L);
    r.d2 = D2[cL];
    r.d3 = D3[cL];
#pragma unroll
    for (int g = 0; g < G; g++) {
      r.l[g] = lam_in[tb[g] + cL];
      r.f[g] = f[tb[g] + cL];
    }
  };
  auto enter = [&](int L, int q, const Raw& r) {
    const bool kdo = L <= kmax && L >= 1 && L <= kmax - kshift;
    CellOp<EX> cw;
    if (kdo) cw = CellOp<EX>::make(r.d2, r.d3, dt);
#pragma unroll
    for (int g = 0; g < G; g++) {
      const TF l = r.l[g];
      w[g][q].f = r.f[g];
      w[g][q].acc = kdo ? cw.diag(l) : l;
      w[g][q].mu = kdo ? cw.mu(l, dt, 1.0) : TF(0);
    }
  };
  // the window at face Fs holds levels Fs-1..Fs+2 (level -1 is zero)
  {
    Raw r0, r1, r2, r3;
    fetch(max(Fs - 1, 0), r0);
    fetch(Fs, r1);
    fetch(Fs + 1, r2);
    fetch(Fs + 2, r3);
    if (Fs >= 1) {
      enter(Fs - 1, 0, r0);
    } else {
#pragma unroll
      for (int g = 0; g < G; g++) w[g][0] = Lev{TF(0), TF(0), TF(0)};
    }
    enter(Fs, 1, r1);
    enter(Fs + 1, 2, r2);
    enter(Fs + 2, 3, r3);
  }
  Raw rq0, rq1;  // levels K'+3, K'+4
  fetch(Fs + 3, rq0);
  fetch(Fs + 4, rq1);
  // D2 at K', K'+1, K'+2 and ww at K', K'+1, K'+2, loaded two faces ahead
  double dq[3], wq[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    dq[q] = D2[lv(Fs + q)];
    wq[q] = ww[lv(Fs + q)];
  }
  for (int Kf = Fs; Kf <= Fe; Kf++) {
    // level 0 is never updated (mu = 0), so at K' = 0 Fbar is -mu(1) with
    // kshift and 0 without
    const bool used = Kf >= 1 || kshift;
    WFace<EX> wf;
    wface_setup<SCH>(wf, wq[0], dt, dq[0], dq[1]);
    const bool lim_up = SCH != UPSTREAM && Kf > 1 - kshift;
    const bool lim_dn = SCH != UPSTREAM && Kf < kmax - 1;
#pragma unroll
    for (int g = 0; g < G; g++) {
      const TF Fbar = used ? w[g][1].mu - w[g][2].mu : TF(0);
      const TF fkm1 = Kf >= 1 ? w[g][0].f : TF(0), fk = w[g][1].f, fk1 = w[g][2].f;
      const TF fk2 = Kf + 2 <= kmax ? w[g][3].f : fk;
      TF gr[4];
      w_flux_grad<TF, SCH>(wf, lim_up, lim_dn, fkm1, fk, fk1, fk2, gr);
      // (a slot past either end of the column is never stored)
#pragma unroll
      for (int q = 0; q < 4; q++) w[g][q].acc += Fbar * gr[q];
    }
    if (Kf - 1 >= L0 && Kf - 1 <= L1) {
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
    enter(Kf + 3, 3, rq0);
    rq0 = rq1;
    fetch(Kf + 5, rq1);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      dq[q] = dq[q + 1];
      wq[q] = wq[q + 1];
    }
    dq[2] = D2[lv(Kf + 3)];
    wq[2] = ww[lv(Kf + 3)];
  }
  // after the last face, levels kmax-1 and kmax are complete
  if (Fe == kmax - 1) {
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (t0 + g < ntr) {
        if (kmax - 1 >= L0 && kmax - 1 <= L1) lam_out[tb[g] + c2 + NXY * (kmax - 1)] = w[g][0].acc;
        if (L1 == kmax) lam_out[tb[g] + c2 + NXY * kmax] = w[g][1].acc;
      }
    }
  }
}

// ---- launchers --------------------------------------------------------------------
hipError_t launch_heights(hipStream_t st, Dims d, const uint8_t* flags, const double* ho,
                          const double* uu, const double* vv, const double* ww,
                          double* const D[5], double dt, double dx, double dy, double ard1,
                          int scheme, int kshift) {
  hipLaunchKernelGGL(k_heights, dim3((d.NX + 63) / 64, (d.NY + 3) / 4, d.kmax + 1), dim3(64, 4),
                     0, st, d, flags, ho, uu, vv, ww, D[0], D[1], D[2], D[3], D[4], dt, dx, dy,
                     ard1, scheme == UPSTREAM ? 0 : 1, kshift);
  return hipGetLastError();
}

// ADJ: a = f, b = lam_in, out = lam_out; else a = fin, out = fout
template <typename TF, bool ADJ, int SCH, bool DIFF>
static void launch_h_one(hipStream_t st, Dims d, int dir, int ntr, long ts, const TF* a,
                         const TF* b, TF* out, const double* Dio, const double* Dn,
                         const double* vel, const double* Dv, const uint8_t* flags, double dt,
                         double dx, double dy, double ard1, double AH) {
  const int ng = (ntr + kGroup<TF> - 1) / kGroup<TF>;
  const dim3 gv((d.NX + 63) / 64, (d.NY + kVRows<TF> - 1) / kVRows<TF>, (d.kmax + 1) * ng);
  if constexpr (ADJ) {
    if (dir == 0)
      hipLaunchKernelGGL((k_sub_u_adj<TF, SCH, DIFF>),
                         dim3((d.NX + kUOutA - 1) / kUOutA, (d.NY + 3) / 4, d.kmax + 1),
                         dim3(64, 4), 0, st, d, ntr, ts, a, b, out, Dio, Dn, vel, Dv, flags, dt,
                         dx, dy, ard1, AH);
    else
      hipLaunchKernelGGL((k_sub_v_adj<TF, SCH, DIFF>), gv, dim3(64), 0, st, d, ntr, ts, a, b, out,
                         Dio, Dn, vel, Dv, flags, dt, dx, dy, ard1, AH, ng);
  } else {
    (void)b;
    if (dir == 0)
      hipLaunchKernelGGL((k_sub_u<TF, SCH, DIFF>),
                         dim3((d.NX + kUOutF - 1) / kUOutF, (d.NY + 3) / 4, d.kmax + 1),
                         dim3(64, 4), 0, st, d, ntr, ts, a, out, Dio, Dn, vel, Dv, flags, dt, dx,
                         dy, ard1, AH);
    else
      hipLaunchKernelGGL((k_sub_v<TF, SCH, DIFF>), gv, dim3(64), 0, st, d, ntr, ts, a, out, Dio,
                         Dn, vel, Dv, flags, dt, dx, dy, ard1, AH, ng);
  }
}

template <typename TF, bool ADJ>
static hipError_t launch_h(hipStream_t st, Dims d, int dir, int ntr, long ts, const TF* a,
                           const TF* b, TF* out, const double* Dio, const double* Dn,
                           const double* vel, const double* Dv, const uint8_t* flags, double dt,
                           double dx, double dy, double ard1, int scheme, double AH) {
#define GT_H(S)                                                                                \
  (AH > 0.0 ? launch_h_one<TF, ADJ, S, true>(st, d, dir, ntr, ts, a, b, out, Dio, Dn, vel, Dv,  \
                                             flags, dt, dx, dy, ard1, AH)                      \
            : launch_h_one<TF, ADJ, S, false>(st, d, dir, ntr, ts, a, b, out, Dio, Dn, vel, Dv, \
                                              flags, dt, dx, dy, ard1, AH))
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

template <typename TF, bool ADJ, int SCH>
static void launch_w_one(hipStream_t st, Dims d, int ntr, long ts, const TF* a, const TF* b,
                         TF* out, const double* D2, const double* D3, const double* ww,
                         const uint8_t* flags, double dt, int kshift) {
  // A column march has one wavefront per 64 columns and tracer group; where
  // that leaves the chip short of wavefronts (one FP64 tracer on a small
  // grid) the columns are cut into up to four chunks of at least 8 levels.
  const int ng = (ntr + kGroup<TF> - 1) / kGroup<TF>;
  const long waves = (long)((d.NX + 63) / 64) * d.NY * ng;
  int nch = 1;
  while (nch < 4 && waves * nch < kWaveTarget && d.kmax / (nch + 1) >= 8) nch++;
  const int nlev = ADJ ? d.kmax + 1 : d.kmax;  // levels split: 0..kmax, resp. 1..kmax
  const int kc = (nlev + nch - 1) / nch;
  const int nchunks = (nlev + kc - 1) / kc;
  const dim3 g((unsigned)(((d.NX + 63) / 64) * ng * nchunks), (d.NY + 3) / 4, 1);
  if constexpr (ADJ) {
    hipLaunchKernelGGL((k_sub_w_adj<TF, SCH>), g, dim3(64, 4), 0, st, d, ntr, ts, a, b, out, D2,
                       D3, ww, flags, dt, kshift, ng, nchunks, kc);
  } else {
    (void)b;
    hipLaunchKernelGGL((k_sub_w<TF, SCH>), g, dim3(64, 4), 0, st, d, ntr, ts, a, out, D2, D3, ww,
                       flags, dt, kshift, ng, nchunks, kc);
  }
}

template <typename TF, bool ADJ>
static hipError_t launch_w(hipStream_t st, Dims d, int ntr, long ts, const TF* a, const TF* b,
                           TF* out, const double* D2, const double* D3, const double* ww,
                           const uint8_t* flags, double dt, int scheme, int kshift) {
  switch (scheme) {
    case UPSTREAM: launch_w_one<TF, ADJ, UPSTREAM>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case P2: launch_w_one<TF, ADJ, P2>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case SUPERBEE: launch_w_one<TF, ADJ, SUPERBEE>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case MUSCL: launch_w_one<TF, ADJ, MUSCL>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    case P2_PDM: launch_w_one<TF, ADJ, P2_PDM>(st, d, ntr, ts, a, b, out, D2, D3, ww, flags, dt, kshift); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename TF>
hipError_t launch_sub_h(hipStream_t st, Dims d, int dir, int ntr, long ts, const TF* fin, TF* fout,
                        const double* Dio, const double* Dn, const double* vel, const double* Dv,
                        const uint8_t* flags, double dt, double dx, double dy, double ard1,
                        int scheme, double AH) {
  return launch_h<TF, false>(st, d, dir, ntr, ts, fin, nullptr, fout, Dio, Dn, vel, Dv, flags, dt,
                             dx, dy, ard1, scheme, AH);
}

template <typename TF>
hipError_t launch_sub_h_adj(hipStream_t st, Dims d, int dir, int ntr, long ts, const TF* f,
                            const TF* lam_in, TF* lam_out, const double* Dio, const double* Dn,
                            const double* vel, const double* Dv, const uint8_t* flags, double dt,
                            double dx, double dy, double ard1, int scheme, double AH) {
  return launch_h<TF, true>(st, d, dir, ntr, ts, f, lam_in, lam_out, Dio, Dn, vel, Dv, flags, dt,
                            dx, dy, ard1, scheme, AH);
}

template <typename TF>
hipError_t launch_sub_w(hipStream_t st, Dims d, int ntr, long ts, const TF* fin, TF* fout,
                        const double* D2, const double* D3, const double* ww,
                        const uint8_t* flags, double dt, int scheme, int kshift) {
  return launch_w<TF, false>(st, d, ntr, ts, fin, nullptr, fout, D2, D3, ww, flags, dt, scheme,
                             kshift);
}

template <typename TF>
hipError_t launch_sub_w_adj(hipStream_t st, Dims d, int ntr, long ts, const TF* f,
                            const TF* lam_in, TF* lam_out, const double* D2, const double* D3,
                            const double* ww, const uint8_t* flags, double dt, int scheme,
                            int kshift) {
  return launch_w<TF, true>(st, d, ntr, ts, f, lam_in, lam_out, D2, D3, ww, flags, dt, scheme,
                            kshift);
}

#define GT_INST(TF)                                                                            \
  template hipError_t launch_sub_h<TF>(hipStream_t, Dims, int, int, long, const TF*, TF*,      \
                                       const double*, const double*, const double*,            \
                                       const double*, const uint8_t*, double, double, double,  \
                                       double, int, double);                                   \
  template hipError_t launch_sub_h_adj<TF>(hipStream_t, Dims, int, int, long, const TF*,       \
                                           const TF*, TF*, const double*, const double*,       \
                                           const double*, const double*, const uint8_t*,       \
                                           double, double, double, double, int, double);       \
  template hipError_t launch_sub_w<TF>(hipStream_t, Dims, int, long, const TF*, TF*,           \
                                       const double*, const double*, const double*,            \
                                       const uint8_t*, double, int, int);                      \
  template hipError_t launch_sub_w_adj<TF>(hipStream_t, Dims, int, long, const TF*, const TF*, \
                                           TF*, const double*, const double*, const double*,   \
                                           const uint8_t*, double, int, int);
GT_INST(float)
GT_INST(double)
#undef GT_INST

}  // namespace gt
