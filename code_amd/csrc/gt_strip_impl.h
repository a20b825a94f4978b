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
//     the cell update (Dio*f - dt*advn)/Di is done in FP64 with 1/Di shared by
//     the tracers.
//
// Which cells each sub-step updates follows the reference exactly (the isub /
// jsub bounds of adv_split_u/v.F90 and the column range of adv_split_w.F90),
// so cells the reference leaves alone pass through unchanged.
#pragma once

#include "gt_batched.h"

namespace gt {

// ---- cross-lane moves (DPP wave shifts; call with every lane active) -------
__device__ __forceinline__ int from_lower_lane(int v) {  // lane l gets lane l-1 (0 in lane 0)
  return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int from_upper_lane(int v) {  // lane l gets lane l+1 (0 in lane 63)
  return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, false);
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
  double vel, cfl, Dv;
  uint8_t code;
};
// fast: vel = 0 for an inactive face, g = 0 where the limiter does not apply
template <>
struct FaceC<false> {
  float vel, g, A, al, E, kA;
  bool right;
};

// face constants from the transport (vel), layer height (Dv) and masks;
// lim_l/lim_r: the flux mask of the upwind neighbour allows the limiter
// (adv_split_u.F90:219-265, adv_split_v.F90:77-119)
template <int SCH>
__device__ __forceinline__ void face_setup(FaceC<true>& fc, bool active, bool lim_l, bool lim_r,
                                           double vel, double Dv, double dti, double dlen,
                                           double /*AH*/) {
  fc.vel = vel;
  fc.Dv = Dv;
  fc.cfl = 0.0;
  if (!active) {
    fc.code = F_NONE;
  } else if (vel > 0.0) {
    const bool lim = SCH != UPSTREAM && lim_l;
    fc.code = lim ? F_LEFT_LIM : F_LEFT;
    if (lim) fc.cfl = vel / Dv * dti / dlen;
  } else {
    const bool lim = SCH != UPSTREAM && lim_r;
    fc.code = lim ? F_RIGHT_LIM : F_RIGHT;
    if (lim) fc.cfl = -vel / Dv * dti / dlen;
  }
}

template <int SCH>
__device__ __forceinline__ void face_setup(FaceC<false>& fc, bool active, bool lim_l, bool lim_r,
                                           double vel, double Dv, double dti, double dlen,
                                           double AH) {
  const bool right = !(vel > 0.0);
  const bool lim = active && SCH != UPSTREAM && (right ? lim_r : lim_l);
  const float c = lim ? (float)(__builtin_fabs(vel) * (dti / dlen) * rcp_nr(Dv)) : 0.0f;
  fc.right = right;
  fc.vel = active ? (float)vel : 0.0f;
  fc.g = lim ? 0.5f * (1.0f - c) : 0.0f;
  fc.A = 2.0f * __builtin_amdgcn_rcpf(c + 1.e-10f);
  fc.al = 0.5f + (1.0f / 6.0f) * (1.0f - 2.0f * c);
  fc.E = 2.0f * __builtin_amdgcn_rcpf(1.0f - c);
  fc.kA = (active && AH > 0.0) ? (float)(AH * Dv / dlen) : 0.0f;
}

// the limiter, division-free FP32 form.  With deltaf = fd-fu,
// deltafu = fu-fuu, r = deltafu/deltaf and g = (1-c)/2 the reference returns
// fu + g*phi(r)*deltaf when deltaf*deltafu > 0 (else fu); phi(r)*deltaf is
// the min (deltaf > 0) or max (deltaf < 0) of the branch terms times deltaf.
// P2-PDM: A = 2/(c+1e-10), al = 1/2 + (1-2c)/6, E = 2/(1-c).
// (adv_interfacial_reconstruction, src/2d/advection.F90:785-879)
template <int SCH>
__device__ __forceinline__ float rec_fast(float fuu, float fu, float fd, float g, float A,
                                          float al, float E) {
  const float df = fd - fu, du = fu - fuu;
  if constexpr (SCH == P2) {
    return fu + g * (al * df + (1.0f - al) * du);
  } else if constexpr (SCH == UPSTREAM) {
    return fu;
  } else {
    float t;
    if constexpr (SCH == P2_PDM) {
      const float t1 = A * du, t2 = du + al * (df - du), t3 = E * df;
      t = df > 0.0f ? fminf(fminf(t1, t2), t3) : fmaxf(fmaxf(t1, t2), t3);
    } else if constexpr (SCH == SUPERBEE) {
      t = df > 0.0f ? fmaxf(fminf(2.0f * du, df), fminf(du, 2.0f * df))
                    : fminf(fmaxf(2.0f * du, df), fmaxf(du, 2.0f * df));
    } else {  // MUSCL
      const float h = 0.5f * (df + du);
      t = df > 0.0f ? fminf(fminf(2.0f * du, h), 2.0f * df)
                    : fmaxf(fmaxf(2.0f * du, h), 2.0f * df);
    }
    return df * du > 0.0f ? fu + g * t : fu;
  }
}

// flux through the face between cells X and X+1 of a sweep, from the cells
// X-1, X, X+1, X+2 (fm1, f0, fp1, fp2)
template <typename TF, int SCH>
__device__ __forceinline__ TF face_flux(const FaceC<true>& fc, TF fm1, TF f0, TF fp1, TF fp2,
                                        double AH, double dlen) {
  if (fc.code == F_NONE) return TF(0);
  TF fu;
  if (fc.code <= F_LEFT_LIM) {
    fu = f0;
    if (fc.code == F_LEFT_LIM) fu = reconstruct<double>(SCH, fc.cfl, fm1, f0, fp1);
  } else {
    fu = fp1;
    if (fc.code == F_RIGHT_LIM) fu = reconstruct<double>(SCH, fc.cfl, fp2, fp1, f0);
  }
  TF fl = fc.vel * fu;
  if (AH > 0.0) fl = fl - AH * fc.Dv * (fp1 - f0) / dlen;
  return fl;
}

template <typename TF, int SCH>
__device__ __forceinline__ TF face_flux(const FaceC<false>& fc, TF fm1, TF f0, TF fp1, TF fp2,
                                        double AH, double /*dlen*/) {
  const TF fuu = fc.right ? fp2 : fm1;
  const TF fu0 = fc.right ? fp1 : f0;
  const TF fd = fc.right ? f0 : fp1;
  const TF fu = rec_fast<SCH>(fuu, fu0, fd, fc.g, fc.A, fc.al, fc.E);
  TF fl = fc.vel * fu;
  if (AH > 0.0) fl = fl - fc.kA * (fp1 - f0);
  return fl;
}

// the tracer update of one cell of a horizontal sub-step.  EXACT: the
// reference's (Dio*fo - dt*advn)/Dn, advn = splitfac*(Fc*dlen - Fm*dlen)*ard1
// (adv_split_u.F90:267-286).  fast: the same in FP64 with the shared
// reciprocal rDn (dtc = dt*splitfac*dlen*ard1).
template <typename TF, bool EXACT>
__device__ __forceinline__ TF cell_update(double Dio, double Dn, double rDn, TF fo, TF Fc, TF Fm,
                                          double dtc, double dlen, double ard1, double dt,
                                          double* advn_out) {
  if constexpr (EXACT) {
    (void)rDn; (void)dtc;
    const double advn = 0.5 * (Fc * dlen - Fm * dlen) * ard1;
    *advn_out = advn;
    return (Dio * fo - dt * advn) / Dn;
  } else {
    (void)Dn; (void)dlen; (void)ard1; (void)dt;
    const float dF = Fc - Fm;
    *advn_out = 0.0;
    return (TF)((Dio * (double)fo - dtc * (double)dF) * rDn);
  }
}

// ---- vertical faces (adv_split_w.F90:249-277) --------------------------------
template <bool EXACT>
struct WFace;
template <>
struct WFace<true> {
  double w, cfl;
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
}
template <int SCH>
__device__ __forceinline__ void wface_setup(WFace<false>& wf, double w, double dt, double Dlo,
                                            double Dhi) {
  const float c = (float)(__builtin_fabs(w) * dt * (2.0 * rcp_nr(Dlo + Dhi)));
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
    if (lim) fu = reconstruct<double>(SCH, wf.cfl * 1.0, fuu, fu, fd);
  } else {
    fu = rec_fast<SCH>(fuu, fu, fd, lim ? wf.g : 0.0f, wf.A, wf.al, wf.E);
  }
  return TF(wf.w) * fu;
}

// ---- work items ------------------------------------------------------------------
struct StripItem {
  int dom, lev, sx, sy, grp;
};

// Blocks of 4 wavefronts, one work item per wavefront.  Consecutive blocks
// are dealt round-robin to the 8 XCDs, so block b is given the work of
// block (b%8)*per + b/8: each XCD takes a contiguous run of items, in which
// the levels of one strip follow each other and meet in that XCD's L2.
template <typename TF>
__device__ __forceinline__ bool strip_item(const FusedArgs<TF>& A, StripItem& it) {
  const long nblk = (A.nitems + 3) / 4;
  const long per = (nblk + 7) / 8;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const long item = ((long)(blockIdx.x % 8) * per + blockIdx.x / 8) * 4 + wave;
  if (item >= A.nitems) return false;
  int di = 0;
  while (di + 1 < A.ndom && A.doms[di + 1].item_begin <= item) di++;
  const FusedDomain<TF>& a = A.doms[di];
  long loc = item - a.item_begin;
  it.dom = di;
  it.lev = 1 + (int)(loc % a.d.kmax);
  loc /= a.d.kmax;
  it.sx = (int)(loc % a.nsx);
  loc /= a.nsx;
  it.sy = (int)(loc % a.nsy);
  it.grp = (int)(loc / a.nsy);
  return true;
}

template <typename T, int N>
__device__ __forceinline__ void push(T (&w)[N], T v) {
#pragma unroll
  for (int i = 0; i + 1 < N; i++) w[i] = w[i + 1];
  w[N - 1] = v;
}

// ---- kernel 1: u/2 then v/2 at level L ------------------------------------------------
template <typename TF, int SCH, int G, bool EXACT, bool ADV>
__global__ __launch_bounds__(256) void k_strip_uv(FusedArgs<TF> A) {
  StripItem it;
  if (!strip_item(A, it)) return;
  const FusedDomain<TF>& a = A.doms[it.dom];
  const int lane = threadIdx.x & 63;
  const int NX = a.d.NX, NY = a.d.NY;
  const long NXY = a.d.NXY;
  const int x = it.sx * kStripOut - 2 + lane;
  const int y0 = it.sy * A.rows;
  const int t0 = it.grp * G;
  const int ntr = min(G, A.ntracers - t0);
  const bool xin = x >= 0 && x < NX;
  const bool xout = xin && lane >= 2 && lane < 2 + kStripOut;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const bool uface_x = x >= isub && x <= NX - 2 - isub;
  const bool uupd_x = x >= 1 + isub && x <= NX - 2 - isub;
  const double dt = A.dt, dth = 0.5 * A.dt, dx = a.dx, dy = a.dy, ard1 = a.ard1, AH = A.AH;
  const double dtcu = dt * (0.5 * dy * ard1), dtcv = dt * (0.5 * dx * ard1);
  const long oL = NXY * (long)it.lev;
  const long ts = a.tstride;
  const TF* fin = a.fin + (long)t0 * ts + oL;
  TF* f2 = a.f2 + (long)t0 * ts + oL;
  const double *ho = a.ho + oL, *uu = a.uu + oL, *hu = a.hu + oL, *vv = a.vv + oL,
               *hv = a.hv + oL;
  double* d2 = a.d2 + oL;
  double* a2 = ADV ? a.a2 + oL : nullptr;
  const int nsteps = A.rows + 4;

  struct Row {
    int fl;
    double ho, uu, hu, vv, hv;
    TF f[G];
  };
  auto load_row = [&](int y, Row& r) {
    const bool in = xin && y >= 0 && y < NY;
    const long c = in ? (long)x + (long)NX * y : 0;
    r.fl = in ? (int)a.flags[c] : 0;
    r.ho = in ? ho[c] : 1.0;
    r.uu = in ? uu[c] : 0.0;
    r.hu = in ? hu[c] : 1.0;
    r.vv = in ? vv[c] : 0.0;
    r.hv = in ? hv[c] : 1.0;
#pragma unroll
    for (int g = 0; g < G; g++) r.f[g] = (in && g < ntr) ? fin[(long)g * ts + c] : TF(0);
  };

  // windows over the last four rows: index 3 = row y, 0 = row y-3
  int flw[4] = {0, 0, 0, 0};
  double vvw[4] = {0, 0, 0, 0}, hvw[4] = {1, 1, 1, 1}, d1w[4] = {1, 1, 1, 1};
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

  Row nxt;
  load_row(y0 - 2, nxt);
  for (int s = 0; s < nsteps; s++) {
    const int y = y0 - 2 + s;
    const Row cur = nxt;
    if (s + 1 < nsteps) load_row(y + 1, nxt);

    // ---- u/2 on row y: shared part
    const int flm = from_lower_lane(cur.fl), flp = from_upper_lane(cur.fl);
    FaceC<EXACT> fu;
    face_setup<SCH>(fu, (cur.fl & MB_UFLUX) && uface_x, flm & MB_UFLUX, flp & MB_UFLUX, cur.uu,
                    cur.hu, dth, dx, AH);
    const double uum = from_lower_lane(cur.uu);
    const bool updu = (cur.fl & MB_UUPD) && uupd_x;
    const double D0 = cur.ho;
    double D1 = D0;
    if (updu) D1 = D0 - dth * (cur.uu * dy - uum * dy) * ard1;
    const double rD1 = EXACT ? 0.0 : rcp_nr(D1);
    push(flw, cur.fl);
    push(vvw, cur.vv);
    push(hvw, cur.hv);
    push(d1w, D1);

    // ---- v/2 on row yv = y-2: shared part (faces from row y0-1 on)
    const int yv = y - 2;
    const bool vface = s >= 3, vrow = s >= 4;
    FaceC<EXACT> fv;
    bool updv = false;
    double D2 = d1w[1], rD2 = 0.0;
    if (vface) {
      face_setup<SCH>(fv, (flw[1] & MB_VFLUX) && yv >= isub && yv <= NY - 2 - isub,
                      flw[0] & MB_VFLUX, flw[2] & MB_VFLUX, vvw[1], hvw[1], dth, dy, AH);
      updv = (flw[1] & MB_VUPD) && yv >= 1 + isub && yv <= NY - 2 - isub;
      if (updv) D2 = d1w[1] - dth * (vvw[1] * dx - vvw[0] * dx) * ard1;
      if constexpr (!EXACT) rD2 = rcp_nr(D2);
    }
    const bool st_ok = vrow && xout && yv < NY;
    const long cv = (long)x + (long)NX * yv;

#pragma unroll
    for (int g = 0; g < G; g++) {
      // u/2 on row y (adv_split_u.F90:219-286)
      const TF f0 = cur.f[g];
      const TF fm1 = from_lower_lane(f0), fp1 = from_upper_lane(f0);
      const TF fp2 = from_upper_lane(fp1);
      const TF Fu = face_flux<TF, SCH>(fu, fm1, f0, fp1, fp2, AH, dx);
      const TF Fum = from_lower_lane(Fu);
      double advu = 0.0;
      TF F1 = f0;
      if (updu) F1 = cell_update<TF, EXACT>(D0, D1, rD1, f0, Fu, Fum, dtcu, dy, ard1, dt, &advu);
      push(f1w[g], F1);
      if constexpr (ADV) {
#pragma unroll
        for (int q = 0; q < 3; q++) a1w[q][g] = a1w[q + 1][g];
        a1w[3][g] = 0.0 + advu;
      }
      // v/2 on row yv (adv_split_v.F90:77-131)
      if (vface) {
        const TF Fv = face_flux<TF, SCH>(fv, f1w[g][0], f1w[g][1], f1w[g][2], f1w[g][3], AH, dy);
        if (vrow) {
          double advv = 0.0;
          TF F2 = f1w[g][1];
          if (updv)
            F2 = cell_update<TF, EXACT>(d1w[1], D2, rD2, f1w[g][1], Fv, fvp[g], dtcv, dx, ard1,
                                        dt, &advv);
          if (st_ok && g < ntr) f2[(long)g * ts + cv] = F2;
          if constexpr (ADV) {
            if (st_ok && g == 0) a2[cv] = a1w[1][0] + advv;
          }
        }
        fvp[g] = Fv;
      }
    }
    if (st_ok && it.grp == 0) d2[cv] = D2;
  }
}

// ---- kernel 2: w, then v/2 and u/2 at level K -------------------------------------------
template <typename TF, int SCH, int G, bool EXACT, bool ADV>
__global__ __launch_bounds__(256) void k_strip_wvu(FusedArgs<TF> A) {
  StripItem it;
  if (!strip_item(A, it)) return;
  const FusedDomain<TF>& a = A.doms[it.dom];
  const int lane = threadIdx.x & 63;
  const int NX = a.d.NX, NY = a.d.NY, kmax = a.d.kmax;
  const long NXY = a.d.NXY;
  const int K = it.lev;
  const int kshift = A.kshift;
  const int x = it.sx * kStripOut - 2 + lane;
  const int y0 = it.sy * A.rows;
  const int t0 = it.grp * G;
  const int ntr = min(G, A.ntracers - t0);
  const bool xin = x >= 0 && x < NX;
  const bool xout = xin && lane >= 2 && lane < 2 + kStripOut;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const bool uface_x = x >= isub && x <= NX - 2 - isub;
  const bool uupd_x = x >= 1 + isub && x <= NX - 2 - isub;
  const bool wcol_x = x <= NX - 2;
  const double dt = A.dt, dth = 0.5 * A.dt, dx = a.dx, dy = a.dy, ard1 = a.ard1, AH = A.AH;
  const double dtcu = dt * (0.5 * dy * ard1), dtcv = dt * (0.5 * dx * ard1);
  const long ts = a.tstride;
  const long oK = NXY * (long)K;
  // the w sub-step's levels: the flux at K+1/2 exists below kmax, the one at
  // K-1/2 above level 1 (or at level 1 with kshift); level K is updated up
  // to kmax-kshift (adv_split_w.F90:249-300)
  const bool wk_on = K <= kmax - 1, wp_on = K > 1 || kshift, kdo = K <= kmax - kshift;
  // sources of levels K-2..K+2 of the tracers after u/2 + v/2 (level 0 is
  // the input), and of the layer heights K-1..K+1
  const TF* fsrc[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const int l = K - 2 + m;
    fsrc[m] = (l < 0 || l > kmax) ? nullptr
              : (l == 0 ? a.fin + (long)t0 * ts : a.f2 + (long)t0 * ts + NXY * (long)l);
  }
  const double* dKm1 = K == 1 ? a.ho : a.d2 + NXY * (long)(K - 1);
  const double* dK = a.d2 + oK;
  const double* dKp1 = K < kmax ? a.d2 + oK + NXY : nullptr;
  const double *wwK = a.ww + oK, *wwKm1 = a.ww + oK - NXY;
  const double *vv = a.vv + oK, *hv = a.hv + oK, *uu = a.uu + oK, *hu = a.hu + oK;
  const double* a2 = ADV ? a.a2 + oK : nullptr;
  TF* fout = a.fout + (long)t0 * ts + oK;
  const int nsteps = A.rows + 4;

  struct Row {
    int fl;
    double wwK, wwKm1, dKm1, dK, dKp1, vv, hv, uu, hu, a2;
    TF f[G][5];
  };
  auto load_row = [&](int y, Row& r) {
    const bool in = xin && y >= 0 && y < NY;
    const long c = in ? (long)x + (long)NX * y : 0;
    r.fl = in ? (int)a.flags[c] : 0;
    r.wwK = in ? wwK[c] : 0.0;
    r.wwKm1 = in ? wwKm1[c] : 0.0;
    r.dKm1 = in ? dKm1[c] : 1.0;
    r.dK = in ? dK[c] : 1.0;
    r.dKp1 = (in && dKp1) ? dKp1[c] : 1.0;
    r.vv = in ? vv[c] : 0.0;
    r.hv = in ? hv[c] : 1.0;
    r.uu = in ? uu[c] : 0.0;
    r.hu = in ? hu[c] : 1.0;
    r.a2 = (ADV && in) ? a2[c] : 0.0;
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
      for (int m = 0; m < 5; m++)
        r.f[g][m] = (in && fsrc[m] && g < ntr) ? fsrc[m][(long)g * ts + c] : TF(0);
  };

  int flw[4] = {0, 0, 0, 0};
  double vvw[4] = {0, 0, 0, 0}, hvw[4] = {1, 1, 1, 1}, d3w[4] = {1, 1, 1, 1};
  double uuw[4] = {0, 0, 0, 0}, huw[4] = {1, 1, 1, 1};
  double a3w[ADV ? 4 : 1][G];
  TF f3w[G][4];
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

  Row nxt;
  load_row(y0 - 2, nxt);
  for (int s = 0; s < nsteps; s++) {
    const int y = y0 - 2 + s;
    const Row cur = nxt;
    if (s + 1 < nsteps) load_row(y + 1, nxt);

    // ---- w on row y: shared part
    const bool wcol = (cur.fl & MB_WCOL) && wcol_x && y <= NY - 2;
    const bool wupd = wcol && kdo;
    const double D2K = cur.dK;
    double D3 = D2K;
    if (wupd) D3 = D2K - dt * (cur.wwK - cur.wwKm1);
    const double rD3 = EXACT ? 0.0 : rcp_nr(D3);
    WFace<EXACT> wfk, wfp;
    wface_setup<SCH>(wfk, cur.wwK, dt, D2K, cur.dKp1);
    wface_setup<SCH>(wfp, cur.wwKm1, dt, cur.dKm1, D2K);
    push(flw, cur.fl);
    push(vvw, cur.vv);
    push(hvw, cur.hv);
    push(uuw, cur.uu);
    push(huw, cur.hu);
    push(d3w, D3);

    // ---- v/2 and u/2 on row yv = y-2: shared part
    const int yv = y - 2;
    const bool vface = s >= 3, vrow = s >= 4;
    FaceC<EXACT> fv, fu;
    bool updv = false, updu = false;
    double D4 = d3w[1], rD4 = 0.0, D5 = d3w[1], rD5 = 0.0;
    if (vface) {
      face_setup<SCH>(fv, (flw[1] & MB_VFLUX) && yv >= isub && yv <= NY - 2 - isub,
                      flw[0] & MB_VFLUX, flw[2] & MB_VFLUX, vvw[1], hvw[1], dth, dy, AH);
      updv = (flw[1] & MB_VUPD) && yv >= 1 + isub && yv <= NY - 2 - isub;
      if (updv) D4 = d3w[1] - dth * (vvw[1] * dx - vvw[0] * dx) * ard1;
      if constexpr (!EXACT) rD4 = rcp_nr(D4);
    }
    if (vrow) {
      const int fl = flw[1];
      const int flm = from_lower_lane(fl), flp = from_upper_lane(fl);
      face_setup<SCH>(fu, (fl & MB_UFLUX) && uface_x, flm & MB_UFLUX, flp & MB_UFLUX, uuw[1],
                      huw[1], dth, dx, AH);
      const double uum = from_lower_lane(uuw[1]);
      updu = (fl & MB_UUPD) && uupd_x;
      D5 = D4;
      if (updu) D5 = D4 - dth * (uuw[1] * dy - uum * dy) * ard1;
      if constexpr (!EXACT) rD5 = rcp_nr(D5);
    }
    const bool st_ok = vrow && xout && yv < NY;
    const long cv = (long)x + (long)NX * yv;

#pragma unroll
    for (int g = 0; g < G; g++) {
      // w on row y (adv_split_w.F90:249-300)
      const TF fKm2 = cur.f[g][0], fKm1 = cur.f[g][1], fK = cur.f[g][2];
      const TF fKp1 = cur.f[g][3], fKp2 = cur.f[g][4];
      TF wprev = TF(0), wk = TF(0);
      if (wcol && wp_on)
        wprev = wflux<TF, SCH, EXACT>(wfp, K - 1, kshift, kmax, K > 1 ? fKm2 : TF(0), fKm1, fK,
                                      fKp1);
      if (wcol && wk_on) wk = wflux<TF, SCH, EXACT>(wfk, K, kshift, kmax, fKm1, fK, fKp1, fKp2);
      TF F3 = fK;
      double advw = 0.0;
      if (wupd) {
        if constexpr (EXACT) {
          advw = 1.0 * (wk - wprev);
          F3 = (D2K * fK - dt * advw) / D3;
        } else {
          F3 = (TF)((D2K * (double)fK - dt * (double)(wk - wprev)) * rD3);
        }
      }
      push(f3w[g], F3);
      if constexpr (ADV) {
#pragma unroll
        for (int q = 0; q < 3; q++) a3w[q][g] = a3w[q + 1][g];
        a3w[3][g] = cur.a2 + advw;
      }
      if (vface) {
        // v/2 on row yv (adv_split_v.F90:77-131)
        const TF Fv = face_flux<TF, SCH>(fv, f3w[g][0], f3w[g][1], f3w[g][2], f3w[g][3], AH, dy);
        if (vrow) {
          double advv = 0.0, advu = 0.0;
          TF F4 = f3w[g][1];
          if (updv)
            F4 = cell_update<TF, EXACT>(d3w[1], D4, rD4, f3w[g][1], Fv, fvp[g], dtcv, dx, ard1,
                                        dt, &advv);
          // u/2 on row yv (adv_split_u.F90:219-286)
          const TF fm1 = from_lower_lane(F4), fp1 = from_upper_lane(F4);
          const TF fp2 = from_upper_lane(fp1);
          const TF Fu = face_flux<TF, SCH>(fu, fm1, F4, fp1, fp2, AH, dx);
          const TF Fum = from_lower_lane(Fu);
          TF F5 = F4;
          if (updu)
            F5 = cell_update<TF, EXACT>(D4, D5, rD5, F4, Fu, Fum, dtcu, dy, ard1, dt, &advu);
          if (st_ok && g < ntr) fout[(long)g * ts + cv] = F5;
          if constexpr (ADV) {
            if (st_ok && g == 0 && a.advres) {
              const double A4 = a3w[1][0] + advv;
              a.advres[oK + cv] = A4 + advu;
            }
          }
        }
        fvp[g] = Fv;
      }
    }
    if (st_ok && it.grp == 0 && a.hires) a.hires[oK + cv] = D5;
  }

  // level 0 passes through (the step updates levels 1..kmax)
  if (K == 1) {
    for (int yv = y0; yv < y0 + A.rows && yv < NY; yv++) {
      if (!xout) continue;
      const long c = (long)x + (long)NX * yv;
      for (int g = 0; g < ntr; g++) a.fout[(long)(t0 + g) * ts + c] = a.fin[(long)(t0 + g) * ts + c];
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
  long items = 0;
  for (int i = 0; i < a.ndom; i++) {
    FusedDomain<TF>& d = a.doms[i];
    d.nsx = (d.d.NX + kStripOut - 1) / kStripOut;
    d.nsy = (d.d.NY + a.rows - 1) / a.rows;
    d.item_begin = items;
    items += (long)d.nsx * d.nsy * d.d.kmax * a.ngroups;
  }
  a.nitems = items;
}

template <typename TF, int SCH, int G, bool EXACT, bool ADV>
hipError_t launch_one(hipStream_t st, const FusedArgs<TF>& A) {
  const long nblk = (A.nitems + 3) / 4;
  const long grid = ((nblk + 7) / 8) * 8;
  hipLaunchKernelGGL((k_strip_uv<TF, SCH, G, EXACT, ADV>), dim3((unsigned)grid), dim3(256), 0,
                     st, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_strip_wvu<TF, SCH, G, EXACT, ADV>), dim3((unsigned)grid), dim3(256), 0,
                     st, A);
  return hipGetLastError();
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
