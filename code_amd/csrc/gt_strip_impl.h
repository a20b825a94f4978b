// gt_strip_impl.h -- the HALFSPLIT transport step as two strip kernels.
//
// The step u/2 + v/2 + w + v/2 + u/2 (src/3d/advection_3d.F90:439-547 of
// getm-model/code) is cut after the first v/2:
//   k_strip_uv   u/2 then v/2 at one level; writes the tracers f2 and the
//                layer heights d2 after them (work fields in HBM);
//   k_strip_wvu  w at one level from f2 at the levels K-2..K+2, then v/2 and
//                u/2; writes the result and hires.
// Neither kernel carries state from level to level, so every (strip, level,
// tracer group) is an independent work item.
//
// A work item is one wavefront.  Lane l owns column x = 60*sx - 2 + l; the
// wavefront marches up a strip of rows, one row per step, so
//   * i-neighbours (the u sub-steps) come from the adjacent lanes through DPP
//     wave shifts, and lanes 2..61 have all four ghost columns the two-cell
//     stencil of a flux and the two fluxes of an update need;
//   * j-neighbours (the v sub-steps) are earlier rows of the march, kept in a
//     four-row register window;
//   * k-neighbours (w) are loads of the neighbouring levels of f2.
// There is no LDS and no barrier.  Work shared by the tracers of a group is
// done once per row: the layer-height trajectory (FP64), and per face the
// upwind direction, whether the limiter applies (masks), the Courant number
// and the limiter constants that depend on it.
//
// Two arithmetic flavours:
//   EXACT (TF=double): the reference's expressions in the reference's order,
//     so results are bit-identical to the per-sub-step kernels and the CPU
//     restatement.
//   fast (TF=float): tracer and flux arithmetic in FP32 with a division-free
//     form of each limiter (phi(r)*deltaf as a min/max of products, the sign
//     of deltaf choosing min or max) and reciprocals for the face constants;
//     the layer heights stay FP64 and the cell update (Dio*f - dt*advn)/Di is
//     f + [f*(Dio-Di)/Di - dF*(dt*c/Di)] with the two factors formed once per
//     cell (Dio-Di in FP64, the rest in FP32, as the Courant numbers) and
//     shared by the tracers.
//
// Which cells each sub-step updates follows the reference exactly (the isub /
// jsub bounds of adv_split_u/v.F90 and the column range of adv_split_w.F90),
// so cells the reference leaves alone pass through unchanged.
#pragma once

#include <type_traits>

#include "gt_batched.h"

namespace gt {

// 1/x to about one FP64 ulp: hardware estimate and one Newton step (fast path)
__device__ __forceinline__ double rcp_nr(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

// ---- one face of a sweep -----------------------------------------------------
enum : uint8_t { F_NONE = 0, F_LEFT = 1, F_LEFT_LIM = 2, F_RIGHT = 3, F_RIGHT_LIM = 4 };

template <bool EXACT>
struct FaceC;
// EXACT: the values the reference's flux expression reads
template <>
struct FaceC<true> {
  double vel, cfl, Dv, E;  // E = 2/(1-cfl), P2-PDM's third bound, formed once
  uint8_t code;
};
// fast: vel = 0 for an inactive face, g = 0 where the limiter does not apply
template <>
struct FaceC<false> {
  float vel, g, A, al, E, kA;
  bool right;
};

// the sweep's constants: dti = splitfac*dt, the face length dlen, and for
// the fast path dti/dlen and AH/dlen (0 without diffusion)
struct FaceK {
  double dti, dlen, cscale, kscale;
};
__device__ __forceinline__ FaceK face_k(double dti, double dlen, double AH) {
  return {dti, dlen, dti / dlen, AH > 0.0 ? AH / dlen : 0.0};
}

// face constants from the transport (vel), layer height (Dv) and masks;
// lim_l/lim_r: the flux mask of the upwind neighbour allows the limiter
// (adv_split_u.F90:219-265, adv_split_v.F90:77-119)
template <int SCH>
__device__ __forceinline__ void face_setup(FaceC<true>& fc, bool active, bool lim_l, bool lim_r,
                                           double vel, double Dv, const FaceK& k) {
  // The exact path is branch-free like the fast one: every operand is formed
  // and the result selected, so the division chains of neighbouring faces and
  // tracers share one basic block and interleave; a lane whose value is not
  // selected may compute inf or NaN, which is never used.  One division
  // chain serves both directions: the reference's vel/Dv*dti/dlen going
  // right, -vel/Dv*dti/dlen going left, with the operand selected first.
  const double dti = k.dti, dlen = k.dlen;
  fc.vel = vel;
  fc.Dv = Dv;
  const bool left = vel > 0.0;
  const bool lim = active && SCH != UPSTREAM && (left ? lim_l : lim_r);
  fc.code = !active ? F_NONE : (left ? (lim ? F_LEFT_LIM : F_LEFT) : (lim ? F_RIGHT_LIM : F_RIGHT));
  const double c = (left ? vel : -vel) / Dv * dti / dlen;
  fc.cfl = lim ? c : 0.0;
  fc.E = SCH == P2_PDM ? 2.0 / (1.0 - fc.cfl) : 0.0;
}

// adv_interfacial_reconstruction (src/2d/advection.F90:806-845), branch-free:
// every operation of reconstruct() in its order, P2-PDM's 2/(1-cfl) taken
// from the face (E) instead of formed per tracer.  Returns the limited value
// and sets take where the reference returns it (deltaf*deltafu > 0 for the
// TVD schemes, always for P2, never for upstream) -- fu otherwise -- so the
// caller folds this choice into its own select of the face's value
template <int SCH>
__device__ __forceinline__ double reconstruct_bf(double cfl, double E, double fuu, double fu,
                                                 double fd, bool& take) {
  const double deltaf = fd - fu;
  const double deltafu = fu - fuu;
  if constexpr (SCH == UPSTREAM) {
    (void)E; (void)cfl;
    take = false;
    return fu;
  } else if constexpr (SCH == P2) {
    (void)E;
    take = true;
    return reconstruct_p2<double>(cfl, fu, deltafu, deltaf);
  } else {
    // The reference's min/max are the hardware's: where the result is
    // selected (deltaf*deltafu > 0) the ratio is positive and, with
    // cfl >= 0, no operand is NaN and no two are zeros of opposite sign, so
    // v_min_f64 / v_max_f64 pick the operand the comparison form picks
    const double ratio = deltafu / deltaf;
    double limiter;
    if constexpr (SCH == P2_PDM) {
      const double one6th = 1.0 / 6.0;
      const double x = one6th * (1.0 - 2.0 * cfl);
      limiter = (0.5 + x) + (0.5 - x) * ratio;
      limiter = __builtin_fmin(__builtin_fmin(2.0 * ratio / (cfl + 1.e-10), limiter), E);
    } else if constexpr (SCH == SUPERBEE) {
      (void)E;
      limiter = __builtin_fmax(__builtin_fmin(2.0 * ratio, 1.0), __builtin_fmin(ratio, 2.0));
    } else {  // MUSCL
      (void)E;
      limiter = __builtin_fmin(__builtin_fmin(2.0 * ratio, 0.5 * (1.0 + ratio)), 2.0);
    }
    take = deltaf * deltafu > 0.0;
    return fu + 0.5 * limiter * (1.0 - cfl) * deltaf;
  }
}

template <int SCH>
__device__ __forceinline__ void face_setup(FaceC<false>& fc, bool active, bool lim_l, bool lim_r,
                                           double vel, double Dv, const FaceK& k) {
  const bool right = !(vel > 0.0);
  const bool lim = active && SCH != UPSTREAM && (right ? lim_r : lim_l);
  // the Courant number feeds FP32 arithmetic only: formed in FP32
  const float c0 =
      __builtin_fabsf((float)vel) * (float)k.cscale * __builtin_amdgcn_rcpf((float)Dv);
  const float c = lim ? c0 : 0.0f;
  fc.right = right;
  fc.vel = active ? (float)vel : 0.0f;
  fc.g = lim ? 0.5f * (1.0f - c) : 0.0f;
  fc.A = 2.0f * __builtin_amdgcn_rcpf(c + 1.e-10f);
  fc.al = 0.5f + (1.0f / 6.0f) * (1.0f - 2.0f * c);
  fc.E = 2.0f * __builtin_amdgcn_rcpf(1.0f - c);
  fc.kA = active ? (float)(k.kscale * Dv) : 0.0f;
}

// the limiter, division-free FP32 form.  With deltaf = fd-fu,
// deltafu = fu-fuu, r = deltafu/deltaf and g = (1-c)/2 the reference returns
// fu + g*phi(r)*deltaf when deltaf*deltafu > 0 (else fu); phi(r)*deltaf is
// the min (deltaf > 0) or max (deltaf < 0) of the branch terms times deltaf.
// P2-PDM: A = 2/(c+1e-10), al = 1/2 + (1-2c)/6, E = 2/(1-c).
// (adv_interfacial_reconstruction, src/2d/advection.F90:785-879)
// On the lanes that take the limited value deltaf and deltafu share a sign,
// and the branch terms of deltaf < 0 are those of |deltaf|, |deltafu|
// negated, each rounding alike, so the max there is minus the min over the
// magnitudes: one min (abs taken as operand modifiers) and the sign folded
// into g give bit for bit the min/max selected by the sign.  No divergent
// branch; fused multiply-adds are written out where they shorten the chain.
template <int SCH>
__device__ __forceinline__ float rec_fast(float fuu, float fu, float fd, float g, float A,
                                          float al, float E) {
  const float df = fd - fu, du = fu - fuu;
  if constexpr (SCH == P2) {
    // al*df + (1-al)*du = du + al*(df-du)
    return __builtin_fmaf(g, __builtin_fmaf(al, df - du, du), fu);
  } else if constexpr (SCH == UPSTREAM) {
    return fu;
  } else {
    const float adf = __builtin_fabsf(df), adu = __builtin_fabsf(du);
    float m;
    if constexpr (SCH == P2_PDM) {
      m = __builtin_fminf(__builtin_fminf(A * adu, __builtin_fmaf(al, adf - adu, adu)),
                          E * adf);
    } else if constexpr (SCH == SUPERBEE) {
      const float a2 = 2.0f * adu, b2 = 2.0f * adf;
      m = __builtin_fmaxf(__builtin_fminf(a2, adf), __builtin_fminf(adu, b2));
    } else {  // MUSCL
      const float a2 = 2.0f * adu, b2 = 2.0f * adf, h = 0.5f * (adf + adu);
      m = __builtin_fminf(__builtin_fminf(a2, h), b2);
    }
    const float gs = df > 0.0f ? g : -g;
    const float r = __builtin_fmaf(gs, m, fu);
    return df * du > 0.0f ? r : fu;
  }
}

// flux through the face between cells X and X+1 of a sweep, from the cells
// X-1, X, X+1, X+2 (fm1, f0, fp1, fp2)
// DIFF: horizontal diffusion on (AH > 0), fixed at compile time
template <typename TF, int SCH, bool DIFF>
__device__ __forceinline__ TF face_flux(const FaceC<true>& fc, TF fm1, TF f0, TF fp1, TF fp2,
                                        double AH, double dlen) {
  // upwind operands selected first, then one reconstruction for both
  // directions (left: fm1, f0, fp1; right: fp2, fp1, f0), selected by the
  // face code (branch-free, see face_setup)
  const bool left = fc.code <= F_LEFT_LIM;
  const TF fuu = left ? fm1 : fp2, fd = left ? fp1 : f0;
  const TF fu0 = left ? f0 : fp1;
  bool take;
  const TF rr = reconstruct_bf<SCH>(fc.cfl, fc.E, fuu, fu0, fd, take);
  const TF fu = ((fc.code == F_LEFT_LIM || fc.code == F_RIGHT_LIM) && take) ? rr : fu0;
  TF fl = fc.vel * fu;
  if constexpr (DIFF) fl = fl - AH * fc.Dv * (fp1 - f0) / dlen;
  return fc.code == F_NONE ? TF(0) : fl;
}

template <typename TF, int SCH, bool DIFF>
__device__ __forceinline__ TF face_flux(const FaceC<false>& fc, TF fm1, TF f0, TF fp1, TF fp2,
                                        double /*AH*/, double /*dlen*/) {
  const TF fuu = fc.right ? fp2 : fm1;
  const TF fu0 = fc.right ? fp1 : f0;
  const TF fd = fc.right ? f0 : fp1;
  const TF fu = rec_fast<SCH>(fuu, fu0, fd, fc.g, fc.A, fc.al, fc.E);
  TF fl = fc.vel * fu;
  if constexpr (DIFF) fl = __builtin_fmaf(-fc.kA, fp1 - f0, fl);
  return fl;
}

// the tracer update of one cell of a horizontal sub-step: the reference's
// (Dio*fo - dt*advn)/Dn, advn = splitfac*(Fc*dlen - Fm*dlen)*ard1
// (adv_split_u.F90:267-286), in the reference's order (EXACT); computed on
// every lane and selected by upd (advn_out is 0 where the cell is not
// updated)
__device__ __forceinline__ double cell_update_exact(bool upd, double Dio, double Dn, double fo,
                                                    double Fc, double Fm, double dlen,
                                                    double ard1, double dt, double* advn_out) {
  const double advn = 0.5 * (Fc * dlen - Fm * dlen) * ard1;
  *advn_out = upd ? advn : 0.0;
  const double fn = (Dio * fo - dt * advn) / Dn;
  return upd ? fn : fo;
}

// fast: the same as fo + [fo*(Dio-Dn)/Dn - (Fc-Fm)*(dt*splitfac*dlen*ard1/Dn)]:
// the two factors are formed once per cell in FP32 from the FP64 layer
// heights (Dio-Dn taken in FP64, then rounded) and shared by the tracers
// (CellK); the bracket is a small correction, so their few FP32 roundings
// are far below the final rounding of fo + correction.
struct CellK {
  float am1, b;
};
__device__ __forceinline__ CellK cell_factors(double Dio, double Dn, double dtc) {
  const float r = __builtin_amdgcn_rcpf((float)Dn);
  return {(float)(Dio - Dn) * r, (float)dtc * r};
}
__device__ __forceinline__ float cell_update_fast(const CellK& k, float fo, float Fc, float Fm) {
  return fo + __builtin_fmaf(fo, k.am1, -(Fc - Fm) * k.b);
}

// ---- vertical faces (adv_split_w.F90:249-277) --------------------------------
template <bool EXACT>
struct WFace;
template <>
struct WFace<true> {
  double w, cfl, E;
};
template <>
struct WFace<false> {
  float w, g, A, al, E;
};

// face k+1/2 of a column: w = ww(k), cfl from the layer heights either side
template <int SCH>
__device__ __forceinline__ void wface_setup(WFace<true>& wf, double w, double dt, double Dlo,
                                            double Dhi) {
  wf.w = w;
  wf.cfl = __builtin_fabs(w) * dt / (0.5 * (Dlo + Dhi));
  wf.E = SCH == P2_PDM ? 2.0 / (1.0 - wf.cfl * 1.0) : 0.0;
}
template <int SCH>
__device__ __forceinline__ void wface_setup(WFace<false>& wf, double w, double dt, double Dlo,
                                            double Dhi) {
  const float c =
      __builtin_fabsf((float)w) * (float)(2.0 * dt) * __builtin_amdgcn_rcpf((float)(Dlo + Dhi));
  wf.w = (float)w;
  wf.g = 0.5f * (1.0f - c);
  wf.A = 2.0f * __builtin_amdgcn_rcpf(c + 1.e-10f);
  wf.al = 0.5f + (1.0f / 6.0f) * (1.0f - 2.0f * c);
  wf.E = 2.0f * __builtin_amdgcn_rcpf(1.0f - c);
}

// flux at face k+1/2 from the levels k-1..k+2 (the limiter applies above
// level 1-kshift going up and below kmax-1 going down)
template <typename TF, int SCH, bool EXACT>
__device__ __forceinline__ TF wflux(const WFace<EXACT>& wf, int k, int kshift, int kmax, TF fkm1,
                                    TF fk, TF fk1, TF fk2) {
  TF fu, fuu = TF(0), fd = TF(0);
  bool lim = false;
  if (wf.w > 0) {
    fu = fk;
    if (SCH != UPSTREAM) lim = k > 1 - kshift;
    if (lim) { fuu = fkm1; fd = fk1; }
  } else {
    fu = fk1;
    if (SCH != UPSTREAM) lim = k < kmax - 1;
    if (lim) { fuu = fk2; fd = fk; }
  }
  if constexpr (EXACT) {
    bool take;
    const TF rr = reconstruct_bf<SCH>(wf.cfl * 1.0, wf.E, fuu, fu, fd, take);
    fu = (lim && take) ? rr : fu;
  } else {
    fu = rec_fast<SCH>(fuu, fu, fd, lim ? wf.g : 0.0f, wf.A, wf.al, wf.E);
  }
  return TF(wf.w) * fu;
}

// fast form of wflux: every operand formed, the limiter switched off by g = 0
// (lim_up / lim_dn: whether it applies going up / down at this face)
template <int SCH>
__device__ __forceinline__ float wflux_fast(const WFace<false>& wf, bool lim_up, bool lim_dn,
                                            float fkm1, float fk, float fk1, float fk2) {
  const bool up = wf.w > 0.0f;
  const float fu = up ? fk : fk1, fuu = up ? fkm1 : fk2, fd = up ? fk1 : fk;
  const bool lim = up ? lim_up : lim_dn;
  return wf.w * rec_fast<SCH>(fuu, fu, fd, lim ? wf.g : 0.0f, wf.A, wf.al, wf.E);
}

template <typename T>
__device__ __forceinline__ T sel(bool c, T a, T b) {
  return c ? a : b;
}

// ---- work items ------------------------------------------------------------------
struct StripItem {
  int dom, lev, sx, sy, grp;
};

// Blocks of 4 wavefronts, one work item per wavefront.  Consecutive blocks
// are dealt round-robin to the 8 XCDs, so block b is given the work of
// block (b%8)*per + b/8: each XCD takes a contiguous run of items.  Items of
// the w kernel are numbered tracer group fastest, then level, then strip, so
// the groups of one strip-level (which read the same layer heights,
// transports and masks) and then the neighbouring levels (which read the same
// tracer levels in the w sub-step) run side by side and meet in that XCD's
// L2.
template <int P, typename TF>
__device__ __forceinline__ bool strip_item(const FusedArgs<TF>& A, StripItem& it) {
  const long nblk = (A.nitems[P] + 3) / 4;
  const long per = (nblk + 7) / 8;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long item = ((long)(blockIdx.x % 8) * per + blockIdx.x / 8) * 4 + wave;
  if (item >= A.nitems[P]) return false;
  int di = 0;
  while (di + 1 < A.ndom && A.doms[di + 1].item_begin[P] <= item) di++;
  const FusedDomain<TF>& a = A.doms[di];
  long loc = item - a.item_begin[P];
  it.dom = di;
  it.grp = (int)(loc % A.ngroups);
  loc /= A.ngroups;
  if constexpr (P == 0) {
    // u/v kernel: no coupling across levels; the strips of one column of
    // strips run side by side instead, so the four ghost rows a strip reads
    // below and above are the rows its neighbours read at the same time
    it.sy = (int)(loc % a.nsy[0]);
    loc /= a.nsy[0];
    it.sx = (int)(loc % a.nsx);
    it.lev = 1 + (int)(loc / a.nsx);
  } else {
    it.lev = 1 + (int)(loc % a.d.kmax);
    loc /= a.d.kmax;
    it.sx = (int)(loc % a.nsx);
    it.sy = (int)(loc / a.nsx);
  }
  return true;
}

template <int R>
using Slot = std::integral_constant<int, R>;

// Per-lane column data, kept in vector registers: xc = x (0 outside the
// array, so an address is always valid), xin = 0xff inside the array (else
// 0, masks the flags), xm = the MB_* bits the column's i-bounds allow
// (adv_split_u.F90 isub bounds; the w column stops at NX-2) plus MB_XOUT on
// the 60 columns a strip writes.
struct ColumnBits {
  int xc, xin, xm;
};
__device__ __forceinline__ ColumnBits column_bits(int x, int lane, int NX, int isub) {
  const bool in = x >= 0 && x < NX;
  int xm = MB_VFLUX | MB_VUPD;
  if (x >= isub && x <= NX - 2 - isub) xm |= MB_UFLUX;
  if (x >= 1 + isub && x <= NX - 2 - isub) xm |= MB_UUPD;
  if (x <= NX - 2) xm |= MB_WCOL;
  if (lane >= 2 && lane < 2 + kStripOut) xm |= MB_XOUT;
  return {in ? x : 0, in ? 0xff : 0, in ? xm : 0};
}

// load / store at a wave-uniform base plus a 32-bit byte offset: the base
// stays in scalar registers and the address costs one vector register (the
// saddr form of global_load/store).  A field of one tracer must stay below
// 4 GiB, which the launch plan checks.
template <typename T>
__device__ __forceinline__ T ldo(const T* base, uint32_t off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <typename T>
__device__ __forceinline__ void sto(T* base, uint32_t off, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}

// The row march is unrolled four times: the four-row windows rotate by
// renaming (logical row q of step s lives in slot (s+q)&3), and the inputs of
// a row are loaded one step ahead into one of two row buffers (FP64) or at
// the start of its step (FP32, below).  Loads read
// valid addresses everywhere (cell 0 outside the array, the group's last
// tracer past its end, level K for a level outside 0..kmax); only the flags
// are masked, since every other value of a cell outside the array, or of a
// missing level, reaches discarded results only.

// FP64 (the exact flavour) loads the inputs of a row one step ahead into the
// other of two row buffers; FP32 loads them where the step starts: without
// the second buffer's registers the FP32 kernels run one more wavefront per
// SIMD (uv 129 -> 101, wvu 229 -> 159 VGPRs at 9 tracers), which hides more
// (batched step 2.80 -> 2.74 ms), while FP64 loses more than it gains
// (advection 0.46 -> 0.49 ms with 140 instead of 182 VGPRs in wvu)
template <typename TF>
constexpr bool kRowAhead = std::is_same<TF, double>::value;
// the w kernel's shared fields of a row (masks, w, heights, the v and u
// transports) with its tracers, two rows ahead (true), or where the row's
// step starts (false): with the tracers alone ahead the FP64 w kernel holds
// 161 instead of 179 VGPRs and runs 3 wavefronts per SIMD instead of 2
// (headline step 0.592 -> 0.587 ms); the u/v kernel keeps both ahead (103
// instead of 113 VGPRs would leave it at 4 wavefronts per SIMD)
template <typename TF>
constexpr bool kWvuSharedAhead = false;

// ---- kernel 1: u/2 then v/2 at level L ------------------------------------------------
template <typename TF, int SCH, int G, bool EXACT, bool ADV, bool DIFF>
__global__ __launch_bounds__(256) void k_strip_uv(FusedArgs<TF> A) {
  StripItem it;
  if (!strip_item<0>(A, it)) return;
  const FusedDomain<TF>& a = A.doms[it.dom];
  const int lane = threadIdx.x & 63;
  const int NX = a.d.NX, NY = a.d.NY;
  const long NXY = a.d.NXY;
  const int x = it.sx * kStripOut - 2 + lane;
  const int y0 = it.sy * A.rows[0];
  const int t0 = it.grp * G;
  const int ntr = min(G, A.ntracers - t0);
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const ColumnBits cb = column_bits(x, lane, NX, isub);
  const bool xout = (cb.xm & MB_XOUT) != 0;
  const double dt = A.dt, dth = 0.5 * A.dt, dx = a.dx, dy = a.dy, ard1 = a.ard1, AH = A.AH;
  const double dtcu = dt * (0.5 * dy * ard1), dtcv = dt * (0.5 * dx * ard1);
  const FaceK ku = face_k(dth, dx, AH), kv = face_k(dth, dy, AH);
  const long oL = NXY * (long)it.lev;
  const long ts = a.tstride;
  const TF* fin_g[G];
  TF* f2_g[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const long tg = (long)(t0 + min(g, ntr - 1)) * ts + oL;
    fin_g[g] = a.fin + tg;
    f2_g[g] = a.f2 + tg;
  }
  const double *ho = a.ho + oL, *uu = a.uu + oL, *hu = a.hu + oL, *vv = a.vv + oL,
               *hv = a.hv + oL;
  double* d2 = a.d2 + oL;
  double* a2 = ADV ? a.a2 + oL : nullptr;
  const uint8_t* flags = a.flags;
  const int nsteps = A.rows[0] + 4;
  auto cell = [&](int y) -> uint32_t {
    return (y >= 0 && y < NY) ? (uint32_t)(cb.xc + NX * y) : (uint32_t)cb.xc;
  };

  // inputs of the step of row y: the u/2 row y and the v faces of rows y-2
  // and y-3
  struct Row {
    int fl;
    double ho, uu, hu, vv, vvm, hv;
    TF f[G];
  };
  auto load_row = [&](int y, Row& r) {
    const uint32_t c = cell(y), c2 = cell(y - 2), c3 = cell(y - 3);
    const uint32_t b = c * (uint32_t)sizeof(TF), b8 = c * 8u;
    r.fl = flags[c];
    r.ho = ldo(ho, b8);
    r.uu = ldo(uu, b8);
    r.hu = ldo(hu, b8);
    r.vv = ldo(vv, c2 * 8u);
    r.vvm = ldo(vv, c3 * 8u);
    r.hv = ldo(hv, c2 * 8u);
#pragma unroll
    for (int g = 0; g < G; g++) r.f[g] = ldo(fin_g[g], b);
  };

  int flw[4] = {0, 0, 0, 0};
  double d1w[4] = {1, 1, 1, 1};
  double a1w[ADV ? 4 : 1][G];
  TF f1w[G][4];
  TF fvp[G];  // v flux through the face below the row being updated
#pragma unroll
  for (int g = 0; g < G; g++) {
    fvp[g] = TF(0);
#pragma unroll
    for (int q = 0; q < 4; q++) f1w[g][q] = TF(0);
  }
  if constexpr (ADV) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int g = 0; g < G; g++) a1w[q][g] = 0.0;
  }

  auto step = [&](auto slot, Row& buf, int s) {
    constexpr int R = decltype(slot)::value;
    constexpr int w0 = R & 3, w1 = (R + 1) & 3, w2 = (R + 2) & 3, w3 = (R + 3) & 3;
    const int y = y0 - 2 + s;
    if constexpr (!kRowAhead<TF>) load_row(y, buf);
    const int fl = (y >= 0 && y < NY) ? (buf.fl & cb.xin) : 0;  // raw masks of the cell
    const int fo = (fl | MB_XOUT) & cb.xm;                      // ... with the column bounds

    // ---- u/2 on row y: shared part
    const int flm = from_lower_lane(fl), flp = from_upper_lane(fl);
    FaceC<EXACT> fu;
    face_setup<SCH>(fu, fo & MB_UFLUX, flm & MB_UFLUX, flp & MB_UFLUX, buf.uu, buf.hu, ku);
    const double uum = from_lower_lane(buf.uu);
    const bool updu = fo & MB_UUPD;
    const double D0 = buf.ho;
    double D1 = D0;
    D1 = sel(updu, D0 - dth * (buf.uu * dy - uum * dy) * ard1, D0);
    CellK k1{};
    if constexpr (!EXACT) k1 = cell_factors(D0, D1, dtcu);
    flw[w3] = fo;
    d1w[w3] = D1;

    // ---- v/2 on row yv = y-2: shared part (faces from row y0-1 on)
    const int yv = y - 2;
    const bool vface = s >= 3, vrow = s >= 4;
    FaceC<EXACT> fv;
    bool updv = false;
    double D2 = d1w[w1];
    CellK k2{};
    if (vface) {
      face_setup<SCH>(fv, (flw[w1] & MB_VFLUX) && yv >= isub && yv <= NY - 2 - isub,
                      flw[w0] & MB_VFLUX, flw[w2] & MB_VFLUX, buf.vv, buf.hv, kv);
      updv = (flw[w1] & MB_VUPD) && yv >= 1 + isub && yv <= NY - 2 - isub;
      D2 = sel(updv, d1w[w1] - dth * (buf.vv * dx - buf.vvm * dx) * ard1, D2);
      if constexpr (!EXACT) k2 = cell_factors(d1w[w1], D2, dtcv);
    }
    const bool st_ok = vrow && (flw[w1] & MB_XOUT) && yv < NY;
    const uint32_t cv = (uint32_t)(x + NX * yv);

#pragma unroll
    for (int g = 0; g < G; g++) {
      // u/2 on row y (adv_split_u.F90:219-286)
      const TF fc = buf.f[g];
      const TF fm1 = from_lower_lane(fc), fp1 = from_upper_lane(fc);
      const TF fp2 = from_upper_lane(fp1);
      const TF Fu = face_flux<TF, SCH, DIFF>(fu, fm1, fc, fp1, fp2, AH, dx);
      const TF Fum = from_lower_lane(Fu);
      double advu = 0.0;
      TF F1 = fc;
      if constexpr (EXACT) {
        F1 = cell_update_exact(updu, D0, D1, fc, Fu, Fum, dy, ard1, dt, &advu);
      } else {
        const TF u = cell_update_fast(k1, fc, Fu, Fum);
        F1 = updu ? u : fc;
      }
      f1w[g][w3] = F1;
      if constexpr (ADV) a1w[w3][g] = 0.0 + advu;
      // v/2 on row yv (adv_split_v.F90:77-131)
      if (vface) {
        const TF Fv =
            face_flux<TF, SCH, DIFF>(fv, f1w[g][w0], f1w[g][w1], f1w[g][w2], f1w[g][w3], AH, dy);
        if (vrow) {
          double advv = 0.0;
          TF F2 = f1w[g][w1];
          if constexpr (EXACT) {
            F2 = cell_update_exact(updv, d1w[w1], D2, F2, Fv, fvp[g], dx, ard1, dt, &advv);
          } else {
            const TF u = cell_update_fast(k2, F2, Fv, fvp[g]);
            F2 = updv ? u : F2;
          }
          if (st_ok && g < ntr) sto(f2_g[g], cv * (uint32_t)sizeof(TF), F2);
          if constexpr (ADV) {
            if (st_ok && g == 0) sto(a2, cv * 8u, a1w[w1][0] + advv);
          }
        }
        fvp[g] = Fv;
      }
    }
    if (st_ok && it.grp == 0) sto(d2, cv * 8u, D2);
    if constexpr (kRowAhead<TF>) load_row(y + 2, buf);
  };

  if constexpr (kRowAhead<TF>) {
    Row ra, rb;
    load_row(y0 - 2, ra);
    load_row(y0 - 1, rb);
    for (int s = 0; s < nsteps; s += 4) {
      step(Slot<0>{}, ra, s);
      step(Slot<1>{}, rb, s + 1);
      step(Slot<2>{}, ra, s + 2);
      step(Slot<3>{}, rb, s + 3);
    }
  } else {
    Row r;
    for (int s = 0; s < nsteps; s += 4) {
      step(Slot<0>{}, r, s);
      step(Slot<1>{}, r, s + 1);
      step(Slot<2>{}, r, s + 2);
      step(Slot<3>{}, r, s + 3);
    }
  }

  // level 0 of the work fields (the second kernel reads levels K-2..K+2 of
  // f2 and K-1..K+1 of d2 from one base): the input tracers and ho
  if (it.lev == 1 && xout) {
    for (int yv = y0; yv < y0 + A.rows[0] && yv < NY; yv++) {
      const uint32_t c = (uint32_t)(x + NX * yv);
      for (int g = 0; g < ntr; g++) {
        const long tg = (long)(t0 + g) * ts;
        a.f2[tg + c] = a.fin[tg + c];
      }
      if (it.grp == 0) a.d2[c] = a.ho[c];
    }
  }
}

// ---- kernel 2: w, then v/2 and u/2 at level K -------------------------------------------
// BUD: some domain of the launch forms its budgets in the step (the
// variant without them carries none of their code or registers)
template <typename TF, int SCH, int G, bool EXACT, bool ADV, bool DIFF, bool BUD>
__global__ __launch_bounds__(256) void k_strip_wvu(FusedArgs<TF> A) {
  StripItem it;
  if (!strip_item<1>(A, it)) return;
  const FusedDomain<TF>& a = A.doms[it.dom];
  const int lane = threadIdx.x & 63;
  const int NX = a.d.NX, NY = a.d.NY, kmax = a.d.kmax;
  const long NXY = a.d.NXY;
  const int K = it.lev;
  const int kshift = A.kshift;
  const int x = it.sx * kStripOut - 2 + lane;
  const int y0 = it.sy * A.rows[1];
  const int t0 = it.grp * G;
  const int ntr = min(G, A.ntracers - t0);
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const ColumnBits cb = column_bits(x, lane, NX, isub);
  const bool xout = (cb.xm & MB_XOUT) != 0;
  const double dt = A.dt, dth = 0.5 * A.dt, dx = a.dx, dy = a.dy, ard1 = a.ard1, AH = A.AH;
  const double dtcu = dt * (0.5 * dy * ard1), dtcv = dt * (0.5 * dx * ard1);
  const FaceK ku = face_k(dth, dx, AH), kv = face_k(dth, dy, AH);
  const long ts = a.tstride;
  const long oK = NXY * (long)K;
  // the w sub-step's levels: the flux at K+1/2 exists below kmax, the one at
  // K-1/2 above level 1 (or at level 1 with kshift); level K is updated up
  // to kmax-kshift (adv_split_w.F90:249-300)
  const bool wk_on = K <= kmax - 1, wp_on = K > 1 || kshift, kdo = K <= kmax - kshift;
  // where the limiter applies at the faces K-1/2 and K+1/2, going up / down
  const bool limp_up = SCH != UPSTREAM && K - 1 > 1 - kshift, limp_dn = SCH != UPSTREAM && K - 1 < kmax - 1;
  const bool limk_up = SCH != UPSTREAM && K > 1 - kshift, limk_dn = SCH != UPSTREAM && K < kmax - 1;
  // tracers after u/2 + v/2: one base per tracer (level 0), levels K-2..K+2
  // as byte offsets; layer heights K-1..K+1 (level 0 of both holds the input)
  const TF* f2_g[G];
  TF* fo_g[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const long tg = (long)(t0 + min(g, ntr - 1)) * ts;
    f2_g[g] = a.f2 + tg;
    fo_g[g] = a.fout + tg + oK;
  }
  uint32_t lvo[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int l = K - 2 + m;
    lvo[m] = (uint32_t)(((l >= 0 && l <= kmax) ? l : K) * NXY * (long)sizeof(TF));
  }
  const bool has_kp1 = K < kmax;
  // d2 at K-1, K, K+1 and ww at K-1, K from one base each (level K-1) and
  // byte offsets
  const double *dKm1 = a.d2 + oK - NXY, *wwKm1 = a.ww + oK - NXY;
  const uint32_t lstep = (uint32_t)(NXY * 8), lup = has_kp1 ? 2u * lstep : lstep;
  const double *vv = a.vv + oK, *hv = a.hv + oK, *uu = a.uu + oK, *hu = a.hu + oK;
  const double* a2 = ADV ? a.a2 + oK : nullptr;
  double* hires = a.hires ? a.hires + oK : nullptr;
  const uint8_t* flags = a.flags;
  const int nsteps = A.rows[1] + 4;
  auto cell = [&](int y) -> uint32_t {
    return (y >= 0 && y < NY) ? (uint32_t)(cb.xc + NX * y) : (uint32_t)cb.xc;
  };

  // inputs of the step of row y: the w row y, and the v and u faces of row
  // y-2 (v also y-3)
  struct Row {
    int fl;
    double wwK, wwKm1, dKm1, dK, dKp1, vv, vvm, hv, uu, hu, a2;
    TF f[G][5];
  };
  // the row's shared fields and its tracers, loaded apart: the tracers run
  // ahead of the arithmetic (kRowAhead), the shared fields only as far as
  // kWvuSharedAhead says
  auto load_shared = [&](int y, Row& r) {
    const uint32_t c = cell(y), c2 = cell(y - 2), c3 = cell(y - 3);
    const uint32_t b8 = c * 8u, b82 = c2 * 8u;
    r.fl = flags[c];
    r.wwK = ldo(wwKm1, b8 + lstep);
    r.wwKm1 = ldo(wwKm1, b8);
    r.dKm1 = ldo(dKm1, b8);
    r.dK = ldo(dKm1, b8 + lstep);
    r.dKp1 = ldo(dKm1, b8 + lup);
    r.vv = ldo(vv, b82);
    r.vvm = ldo(vv, c3 * 8u);
    r.hv = ldo(hv, b82);
    r.uu = ldo(uu, b82);
    r.hu = ldo(hu, b82);
    if constexpr (ADV) r.a2 = ldo(a2, b8);
  };
  auto load_f = [&](int y, Row& r) {
    const uint32_t b = cell(y) * (uint32_t)sizeof(TF);
    uint32_t lb[5];
#pragma unroll
    for (int m = 0; m < 5; m++) lb[m] = b + lvo[m];
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
      for (int m = 0; m < 5; m++) r.f[g][m] = ldo(f2_g[g], lb[m]);
  };
  auto load_row = [&](int y, Row& r) {
    load_shared(y, r);
    load_f(y, r);
  };

  int flw[4] = {0, 0, 0, 0}, fow[4] = {0, 0, 0, 0};
  double d3w[4] = {1, 1, 1, 1};
  double a3w[ADV ? 4 : 1][G];
  TF f3w[G][4];
  // budget (when a.bpart is set): the lane's sums over the rows it writes of
  // hires*f per tracer and of hires, interior water cells only, kept in LDS
  // (one slot per lane, output and wavefront) so they cost no registers
  const bool bud = BUD && a.bpart != nullptr;
  __shared__ double bsh[4][G + 1][64];
  const int wv = threadIdx.x >> 6;
  if (bud) {
#pragma unroll
    for (int q = 0; q <= G; q++) bsh[wv][q][lane] = 0.0;
  }
  TF fvp[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    fvp[g] = TF(0);
#pragma unroll
    for (int q = 0; q < 4; q++) f3w[g][q] = TF(0);
  }
  if constexpr (ADV) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int g = 0; g < G; g++) a3w[q][g] = 0.0;
  }

  auto step = [&](auto slot, Row& buf, int s) {
    constexpr int R = decltype(slot)::value;
    constexpr int w0 = R & 3, w1 = (R + 1) & 3, w2 = (R + 2) & 3, w3 = (R + 3) & 3;
    const int y = y0 - 2 + s;
    if constexpr (!kRowAhead<TF>) {
      load_row(y, buf);
    } else if constexpr (!kWvuSharedAhead<TF>) {
      load_shared(y, buf);
    }
    const int fl = (y >= 0 && y < NY) ? (buf.fl & cb.xin) : 0;  // raw masks of the cell
    const int fo = (fl | MB_XOUT) & cb.xm;                      // ... with the column bounds

    // ---- w on row y: shared part
    const bool wcol = (fo & MB_WCOL) && y <= NY - 2;
    const bool wupd = wcol && kdo;
    const double D2K = buf.dK;
    double D3 = D2K;
    D3 = sel(wupd, D2K - dt * (buf.wwK - buf.wwKm1), D2K);
    CellK k3{};
    if constexpr (!EXACT) k3 = cell_factors(D2K, D3, dt);
    WFace<EXACT> wfk, wfp;
    wface_setup<SCH>(wfk, buf.wwK, dt, D2K, buf.dKp1);
    wface_setup<SCH>(wfp, buf.wwKm1, dt, buf.dKm1, D2K);
    flw[w3] = fl;
    fow[w3] = fo;
    d3w[w3] = D3;

    // ---- v/2 and u/2 on row yv = y-2: shared part
    const int yv = y - 2;
    const bool vface = s >= 3, vrow = s >= 4;
    FaceC<EXACT> fv, fu;
    bool updv = false, updu = false;
    double D4 = d3w[w1], D5 = d3w[w1];
    CellK k4{}, k5{};
    if (vface) {
      face_setup<SCH>(fv, (flw[w1] & MB_VFLUX) && yv >= isub && yv <= NY - 2 - isub,
                      flw[w0] & MB_VFLUX, flw[w2] & MB_VFLUX, buf.vv, buf.hv, kv);
      updv = (flw[w1] & MB_VUPD) && yv >= 1 + isub && yv <= NY - 2 - isub;
      D4 = sel(updv, d3w[w1] - dth * (buf.vv * dx - buf.vvm * dx) * ard1, D4);
      if constexpr (!EXACT) k4 = cell_factors(d3w[w1], D4, dtcv);
    }
    if (vrow) {
      const int flv = flw[w1], fov = fow[w1];
      const int flm = from_lower_lane(flv), flp = from_upper_lane(flv);
      face_setup<SCH>(fu, fov & MB_UFLUX, flm & MB_UFLUX, flp & MB_UFLUX, buf.uu, buf.hu, ku);
      const double uum = from_lower_lane(buf.uu);
      updu = fov & MB_UUPD;
      D5 = D4;
      D5 = sel(updu, D4 - dth * (buf.uu * dy - uum * dy) * ard1, D4);
      if constexpr (!EXACT) k5 = cell_factors(D4, D5, dtcu);
    }
    const bool st_ok = vrow && (fow[w1] & MB_XOUT) && yv < NY;
    const uint32_t cv = (uint32_t)(x + NX * yv);
    // interior water cell of the mass and volume budgets
    const bool bcell = bud && st_ok && (flw[w1] & MB_WCOL) && x >= 2 && x <= NX - 3 &&
                       yv >= 2 && yv <= NY - 3;

#pragma unroll
    for (int g = 0; g < G; g++) {
      // w on row y (adv_split_w.F90:249-300); the level below 1 is zero in
      // the argument, level kmax+1 is never read by the limiter
      const TF fKm2 = buf.f[g][0], fKm1 = buf.f[g][1], fK = buf.f[g][2];
      const TF fKp1 = buf.f[g][3], fKp2 = buf.f[g][4];
      TF wprev = TF(0), wk = TF(0);
      if constexpr (EXACT) {
        const TF p =
            wflux<TF, SCH, EXACT>(wfp, K - 1, kshift, kmax, K > 1 ? fKm2 : TF(0), fKm1, fK, fKp1);
        const TF q = wflux<TF, SCH, EXACT>(wfk, K, kshift, kmax, fKm1, fK, fKp1, fKp2);
        wprev = (wcol && wp_on) ? p : TF(0);
        wk = (wcol && wk_on) ? q : TF(0);
      } else {
        const TF p = wflux_fast<SCH>(wfp, limp_up, limp_dn, K > 1 ? fKm2 : TF(0), fKm1, fK, fKp1);
        const TF q = wflux_fast<SCH>(wfk, limk_up, limk_dn, fKm1, fK, fKp1, fKp2);
        wprev = (wcol && wp_on) ? p : TF(0);
        wk = (wcol && wk_on) ? q : TF(0);
      }
      TF F3 = fK;
      double advw = 0.0;
      if constexpr (EXACT) {
        const double aw = 1.0 * (wk - wprev);
        const double f3 = (D2K * fK - dt * aw) / D3;
        advw = wupd ? aw : 0.0;
        F3 = wupd ? f3 : fK;
      } else {
        const TF u = cell_update_fast(k3, fK, wk, wprev);
        F3 = wupd ? u : fK;
      }
      f3w[g][w3] = F3;
      if constexpr (ADV) a3w[w3][g] = buf.a2 + advw;
      if (vface) {
        // v/2 on row yv (adv_split_v.F90:77-131)
        const TF Fv =
            face_flux<TF, SCH, DIFF>(fv, f3w[g][w0], f3w[g][w1], f3w[g][w2], f3w[g][w3], AH, dy);
        if (vrow) {
          double advv = 0.0, advu = 0.0;
          TF F4 = f3w[g][w1];
          if constexpr (EXACT) {
            F4 = cell_update_exact(updv, d3w[w1], D4, F4, Fv, fvp[g], dx, ard1, dt, &advv);
          } else {
            const TF u = cell_update_fast(k4, F4, Fv, fvp[g]);
            F4 = updv ? u : F4;
          }
          // u/2 on row yv (adv_split_u.F90:219-286)
          const TF fm1 = from_lower_lane(F4), fp1 = from_upper_lane(F4);
          const TF fp2 = from_upper_lane(fp1);
          const TF Fu = face_flux<TF, SCH, DIFF>(fu, fm1, F4, fp1, fp2, AH, dx);
          const TF Fum = from_lower_lane(Fu);
          TF F5 = F4;
          if constexpr (EXACT) {
            F5 = cell_update_exact(updu, D4, D5, F4, Fu, Fum, dy, ard1, dt, &advu);
          } else {
            const TF u = cell_update_fast(k5, F4, Fu, Fum);
            F5 = updu ? u : F4;
          }
          if (st_ok && g < ntr) sto(fo_g[g], cv * (uint32_t)sizeof(TF), F5);
          if (bcell) bsh[wv][g][lane] = __builtin_fma(D5, (double)F5, bsh[wv][g][lane]);
          if constexpr (ADV) {
            if (st_ok && g == 0 && a.advres) {
              const double A4 = a3w[w1][0] + advv;
              a.advres[oK + cv] = A4 + advu;
            }
          }
        }
        fvp[g] = Fv;
      }
    }
    if (st_ok && it.grp == 0 && hires) sto(hires, cv * 8u, D5);
    if (bcell && it.grp == 0) bsh[wv][G][lane] += D5;
    if constexpr (kRowAhead<TF>) {
      if constexpr (kWvuSharedAhead<TF>) {
        load_row(y + 2, buf);
      } else {
        load_f(y + 2, buf);
      }
    }
  };

  if constexpr (kRowAhead<TF>) {
    Row ra, rb;
    if constexpr (kWvuSharedAhead<TF>) {
      load_row(y0 - 2, ra);
      load_row(y0 - 1, rb);
    } else {
      load_f(y0 - 2, ra);
      load_f(y0 - 1, rb);
    }
    for (int s = 0; s < nsteps; s += 4) {
      step(Slot<0>{}, ra, s);
      step(Slot<1>{}, rb, s + 1);
      step(Slot<2>{}, ra, s + 2);
      step(Slot<3>{}, rb, s + 3);
    }
  } else {
    Row r;
    for (int s = 0; s < nsteps; s += 4) {
      step(Slot<0>{}, r, s);
      step(Slot<1>{}, r, s + 1);
      step(Slot<2>{}, r, s + 2);
      step(Slot<3>{}, r, s + 3);
    }
  }

  // the budget partials of this strip-level: the lanes' sums added over the
  // wavefront in a six-level FP64 tree (a partial is then within R + 6
  // roundings of exact for terms of one sign, R the strip's rows), written
  // by lane 0 as a (value, 0) pair; the final pass adds the pairs as
  // double-doubles
  if (bud) {
    const long nsl = (long)a.nsx * a.nsy[1] * kmax;
    const long sl = ((long)(K - 1) * a.nsy[1] + it.sy) * a.nsx + it.sx;
    auto wave_sum = [&](double v, int q) {
      double h = v;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) h += __shfl_xor(h, m);
      if (lane == 0) {
        a.bpart[2 * ((long)q * nsl + sl)] = h;
        a.bpart[2 * ((long)q * nsl + sl) + 1] = 0.0;
      }
    };
#pragma unroll
    for (int g = 0; g < G; g++)
      if (g < ntr) wave_sum(bsh[wv][g][lane], t0 + g);
    if (it.grp == 0) wave_sum(bsh[wv][G][lane], A.ntracers);
  }

  // level 0 passes through (the step updates levels 1..kmax)
  if (K == 1 && xout) {
    for (int yv = y0; yv < y0 + A.rows[1] && yv < NY; yv++) {
      const uint32_t c = (uint32_t)(x + NX * yv);
      for (int g = 0; g < ntr; g++) {
        const long tg = (long)(t0 + g) * ts;
        a.fout[tg + c] = a.fin[tg + c];
      }
      if (it.grp == 0 && a.hires) a.hires[c] = a.ho[c];
      if constexpr (ADV) {
        if (a.advres) a.advres[c] = 0.0;
      }
    }
  }
}

template <typename TF>
void strip_plan(FusedArgs<TF>& a, int group) {
  a.ngroups = (a.ntracers + group - 1) / group;
  for (int p = 0; p < 2; p++) {
    long items = 0;
    for (int i = 0; i < a.ndom; i++) {
      FusedDomain<TF>& d = a.doms[i];
      d.nsx = (d.d.NX + kStripOut - 1) / kStripOut;
      d.nsy[p] = (d.d.NY + a.rows[p] - 1) / a.rows[p];
      d.item_begin[p] = items;
      items += (long)d.nsx * d.nsy[p] * d.d.kmax * a.ngroups;
    }
    a.nitems[p] = items;
  }
}

template <typename TF, int SCH, int G, bool EXACT, bool ADV, bool DIFF>
hipError_t launch_two(hipStream_t st, const FusedArgs<TF>& A) {
  auto grid = [&](int p) { return (unsigned)((((A.nitems[p] + 3) / 4 + 7) / 8) * 8); };
  hipLaunchKernelGGL((k_strip_uv<TF, SCH, G, EXACT, ADV, DIFF>), dim3(grid(0)), dim3(256), 0, st,
                     A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  bool bud = false;
  for (int i = 0; i < A.ndom; i++) bud = bud || A.doms[i].bpart != nullptr;
  if (bud)
    hipLaunchKernelGGL((k_strip_wvu<TF, SCH, G, EXACT, ADV, DIFF, true>), dim3(grid(1)),
                       dim3(256), 0, st, A);
  else
    hipLaunchKernelGGL((k_strip_wvu<TF, SCH, G, EXACT, ADV, DIFF, false>), dim3(grid(1)),
                       dim3(256), 0, st, A);
  return hipGetLastError();
}

template <typename TF, int SCH, int G, bool EXACT, bool ADV>
hipError_t launch_one(hipStream_t st, const FusedArgs<TF>& A) {
  return A.AH > 0.0 ? launch_two<TF, SCH, G, EXACT, ADV, true>(st, A)
                    : launch_two<TF, SCH, G, EXACT, ADV, false>(st, A);
}

// one scheme at compile time (hscheme == vscheme)
template <typename TF, int G, bool EXACT, bool ADV>
hipError_t dispatch_scheme(hipStream_t st, const FusedArgs<TF>& a) {
  if (a.hscheme != a.vscheme) return hipErrorInvalidValue;
  switch (a.hscheme) {
    case P2_PDM: return launch_one<TF, P2_PDM, G, EXACT, ADV>(st, a);
    case SUPERBEE: return launch_one<TF, SUPERBEE, G, EXACT, ADV>(st, a);
    case MUSCL: return launch_one<TF, MUSCL, G, EXACT, ADV>(st, a);
    case P2: return launch_one<TF, P2, G, EXACT, ADV>(st, a);
    case UPSTREAM: return launch_one<TF, UPSTREAM, G, EXACT, ADV>(st, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gt
