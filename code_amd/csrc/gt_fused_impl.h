// gt_fused_impl.h -- the fused HALFSPLIT transport kernel (template).
//
// One pass over HBM for the whole step u/2 + v/2 + w + v/2 + u/2
// (src/3d/advection_3d.F90:439-547 of getm-model/code).
//
// Each workgroup owns a TX x TY tile of water columns and marches up the
// column.  At march iteration m it
//   * runs the two leading horizontal sub-steps (u/2, v/2) on level L = m,
//     with f staged in LDS on the tile plus a 4-cell ghost ring (the stencil
//     radius of two consecutive 2-cell sub-steps);
//   * runs w, then the trailing v/2 and u/2 on level K = m-2, and writes the
//     result: the w sub-step needs the post-v/2 values of levels K-1..K+2,
//     which each thread keeps for its own columns in a 4-deep register ring.
// Work shared by the tracers of a block is done once per level:
//   * the layer-height trajectory hi (FP64) through the five sub-steps;
//   * per face, the upwind direction, whether the limiter applies (masks),
//     the Courant number and the limiter constants that depend on it.
// So each extra tracer costs its own f traffic (read once, written once)
// and its flux arithmetic only.  The march is software-pipelined: the global
// data of the next level is loaded into registers while the current level
// computes.  Every phase maps element l of its region to thread l mod NT, so
// what a thread computed for its own elements stays in its registers.
//
// Two arithmetic flavours:
//   EXACT (TF=double): the reference's expressions in the reference's order
//     (the precomputed Courant number is the same expression), so results are
//     bit-identical to the per-sub-step kernels and the CPU restatement.
//   fast (TF=float): tracer and flux arithmetic in FP32 with a division-free
//     form of each limiter (phi(r)*deltaf evaluated as a min/max of products,
//     the sign of deltaf choosing min or max); the cell update
//     (Dio*f - dt*advn)/Di is done in FP64 with 1/Di shared by all tracers.
//
// Which cells each sub-step updates follows the reference exactly (the isub /
// jsub bounds of adv_split_u/v.F90 and the column range of adv_split_w.F90),
// so cells the reference leaves alone pass through unchanged.
#pragma once

#include "gt_batched.h"

namespace gt {

struct Region {
  int ox, oy, w, h;
  __device__ __forceinline__ int idx(int x, int y) const { return (x - ox) + w * (y - oy); }
};

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// face codes
enum : uint8_t { F_NONE = 0, F_LEFT = 1, F_LEFT_LIM = 2, F_RIGHT = 3, F_RIGHT_LIM = 4 };

template <bool EXACT>
struct FaceBuf;
// EXACT: the face values the reference's flux expression reads
template <>
struct FaceBuf<true> {
  double vel, cfl, Dv;
};
// fast: everything the division-free limiter needs
template <>
struct FaceBuf<false> {
  float vel, g, A, al, E, kA;
};

// the limiter in the reference's form, scheme fixed at compile time
template <int SCH>
__device__ __forceinline__ double rec_exact(double cfl, double fuu, double fu, double fd) {
  return reconstruct<double>(SCH, cfl, fuu, fu, fd);
}

// the limiter, division-free FP32 form.  With deltaf = fd-fu,
// deltafu = fu-fuu, r = deltafu/deltaf and g = (1-c)/2 the reference returns
// fu + g*phi(r)*deltaf when deltaf*deltafu > 0 (else fu); phi(r)*deltaf is
// the min (deltaf > 0) or max (deltaf < 0) of the branch terms times deltaf.
// P2-PDM: A = 2/(c+1e-10), al = 1/2 + (1-2c)/6, E = 2/(1-c).
template <int SCH>
__device__ __forceinline__ float rec_fast(float fuu, float fu, float fd, float g, float A, float al,
                                          float E) {
  const float df = fd - fu, du = fu - fuu;
  if constexpr (SCH == P2) {
    return fu + g * (al * df + (1.0f - al) * du);
  } else if constexpr (SCH == UPSTREAM) {
    return fu;
  } else {
    float t;
    if constexpr (SCH == P2_PDM) {
      float t1 = A * du, t2 = du + al * (df - du), t3 = E * df;
      t = df > 0.0f ? fminf(fminf(t1, t2), t3) : fmaxf(fmaxf(t1, t2), t3);
    } else if constexpr (SCH == SUPERBEE) {
      t = df > 0.0f ? fmaxf(fminf(2.0f * du, df), fminf(du, 2.0f * df))
                    : fminf(fmaxf(2.0f * du, df), fmaxf(du, 2.0f * df));
    } else {  // MUSCL
      float h = 0.5f * (df + du);
      t = df > 0.0f ? fminf(fminf(2.0f * du, h), 2.0f * df)
                    : fmaxf(fmaxf(2.0f * du, h), 2.0f * df);
    }
    return df * du > 0.0f ? fu + g * t : fu;
  }
}

// per-face constants from the face's transport (vel), layer height (Dv) and
// masks; lim_l/lim_r: whether mask_flux at the upwind neighbour allows the
// limiter (adv_split_u.F90:219-265)
template <int SCH, bool EXACT>
__device__ __forceinline__ uint8_t face_setup(FaceBuf<EXACT>& fb, bool active, bool lim_l,
                                              bool lim_r, double vel, double Dv, double dti,
                                              double dlen, double AH) {
  if (!active) {
    fb = {};
    return F_NONE;
  }
  uint8_t code;
  double cfl = 0.0;
  if (vel > 0.0) {
    bool lim = SCH != UPSTREAM && lim_l;
    code = lim ? F_LEFT_LIM : F_LEFT;
    if (lim) cfl = vel / Dv * dti / dlen;
  } else {
    bool lim = SCH != UPSTREAM && lim_r;
    code = lim ? F_RIGHT_LIM : F_RIGHT;
    if (lim) cfl = -vel / Dv * dti / dlen;
  }
  if constexpr (EXACT) {
    fb.vel = vel;
    fb.cfl = cfl;
    fb.Dv = Dv;
  } else {
    const double one6th = 1.0 / 6;
    double x = one6th * (1.0 - 2.0 * cfl);
    fb.vel = (float)vel;
    fb.g = (float)(0.5 * (1.0 - cfl));
    fb.A = (float)(2.0 / (cfl + 1.e-10));
    fb.al = (float)(0.5 + x);
    fb.E = (float)(2.0 / (1.0 - cfl));
    fb.kA = AH > 0.0 ? (float)(AH * Dv / dlen) : 0.0f;
  }
  return code;
}

// flux at a face from its code and constants; f holds the cells along the
// sweep at l0 (cell X) and l0 + n*st
template <typename TF, int SCH, bool EXACT>
__device__ __forceinline__ TF face_flux(uint8_t code, const FaceBuf<EXACT>& fb, const TF* f,
                                        int l0, int st, double AH, double dlen) {
  if (code == F_NONE) return TF(0);
  TF fu;
  const bool left = code <= F_LEFT_LIM;
  const bool lim = code == F_LEFT_LIM || code == F_RIGHT_LIM;
  if constexpr (EXACT) {
    if (left) {
      fu = f[l0];
      if (lim) fu = rec_exact<SCH>(fb.cfl, f[l0 - st], fu, f[l0 + st]);
    } else {
      fu = f[l0 + st];
      if (lim) fu = rec_exact<SCH>(fb.cfl, f[l0 + 2 * st], fu, f[l0]);
    }
    TF fl = fb.vel * fu;
    if (AH > 0.0) fl = fl - AH * fb.Dv * (f[l0 + st] - f[l0]) / dlen;
    return fl;
  } else {
    (void)dlen;
    if (left) {
      fu = f[l0];
      if (lim) fu = rec_fast<SCH>(f[l0 - st], fu, f[l0 + st], fb.g, fb.A, fb.al, fb.E);
    } else {
      fu = f[l0 + st];
      if (lim) fu = rec_fast<SCH>(f[l0 + 2 * st], fu, f[l0], fb.g, fb.A, fb.al, fb.E);
    }
    TF fl = fb.vel * fu;
    if (AH > 0.0) fl = fl - fb.kA * (f[l0 + st] - f[l0]);
    return fl;
  }
}

// the tracer update of one cell of a horizontal sub-step.  EXACT: the
// reference's (Dio*fo - dt*advn)/Dn, advn = splitfac*(Fc*dlen - Fm*dlen)*ard1.
// fast: the same in FP64 with the shared reciprocal rDn (cdl =
// splitfac*dlen*ard1).
template <typename TF, bool EXACT>
__device__ __forceinline__ TF cell_update(double Dio, double Dn, double rDn, TF fo, TF Fc, TF Fm,
                                          double cdl, double dlen, double ard1, double dt,
                                          double* advn_out) {
  if constexpr (EXACT) {
    (void)rDn; (void)cdl;
    double advn = 0.5 * (Fc * dlen - Fm * dlen) * ard1;
    *advn_out = advn;
    return (Dio * fo - dt * advn) / Dn;
  } else {
    (void)Dn; (void)dlen; (void)ard1;
    float dF = Fc - Fm;
    *advn_out = 0.0;
    return (TF)((Dio * (double)fo - (dt * cdl) * (double)dF) * rDn);
  }
}

// vertical flux at level k of a column (adv_split_w.F90:249-277)
template <typename TF, int SCH, bool EXACT>
__device__ __forceinline__ TF wflux(int k, int kshift, int kmax, double w, TF fkm1, TF fk, TF fk1,
                                    TF fk2, double cfl) {
  TF fu, fuu = TF(0), fd = TF(0);
  bool lim = false;
  if (w > 0.0) {
    fu = fk;
    if (SCH != UPSTREAM) lim = k > 1 - kshift;
    if (lim) { fuu = fkm1; fd = fk1; }
  } else {
    fu = fk1;
    if (SCH != UPSTREAM) lim = k < kmax - 1;
    if (lim) { fuu = fk2; fd = fk; }
  }
  if (lim) {
    if constexpr (EXACT) {
      fu = rec_exact<SCH>(cfl * 1.0, fuu, fu, fd);
    } else {
      const double one6th = 1.0 / 6;
      double x = one6th * (1.0 - 2.0 * cfl);
      fu = rec_fast<SCH>(fuu, fu, fd, (float)(0.5 * (1.0 - cfl)), (float)(2.0 / (cfl + 1.e-10)),
                         (float)(0.5 + x), (float)(2.0 / (1.0 - cfl)));
    }
  }
  return TF(w) * fu;
}

template <typename TF, int SCH, int G, bool EXACT, bool ADV, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_fused_halfsplit(FusedArgs<TF> A) {
  using FB = FaceBuf<EXACT>;
  // ---- block -> (domain, tile, tracer group), XCD-aware: blocks
  // b, b+8, b+16, ... run on one XCD, so they get consecutive work items and
  // neighbouring tiles (which share ghost cells) meet in that XCD's L2
  const long nb = A.ntiles * A.ngroups;
  const long per = (nb + 7) / 8;
  const long q0 = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (q0 >= nb) return;
  const long gtile = q0 / A.ngroups;
  const int grp = (int)(q0 % A.ngroups);
  int di = 0;
  while (di + 1 < A.ndom && A.doms[di + 1].tile_begin <= gtile) di++;
  const FusedDomain<TF>& a = A.doms[di];
  const int tile = (int)(gtile - a.tile_begin);
  const int gx0 = (tile % a.ntx) * TX, gy0 = (tile / a.ntx) * TY;
  const int NX = a.d.NX, NY = a.d.NY, kmax = a.d.kmax;
  const long NXY = a.d.NXY;
  const int tid = threadIdx.x;
  const int kshift = A.kshift;
  constexpr int isub = SCH == UPSTREAM ? 0 : 1;
  const double dt = A.dt, dth = 0.5 * A.dt, dx = a.dx, dy = a.dy, ard1 = a.ard1, AH = A.AH;
  const int t0 = grp * G;
  const int ntr = min(G, A.ntracers - t0);
  // fast path: splitfac*dlen*ard1 for u (dlen=dy) and v (dlen=dx)
  const double cdu = 0.5 * dy * ard1, cdv = 0.5 * dx * ard1;

  const Region R0{gx0 - 4, gy0 - 4, TX + 8, TY + 8};
  const Region R1{gx0 - 2, gy0 - 4, TX + 4, TY + 8};
  const Region R2{gx0 - 2, gy0 - 2, TX + 4, TY + 4};
  const Region R4{gx0 - 2, gy0, TX + 4, TY};
  const Region RT{gx0, gy0, TX, TY};
  const Region FU1{gx0 - 3, gy0 - 4, TX + 5, TY + 8};
  const Region FV2{gx0 - 2, gy0 - 3, TX + 4, TY + 5};
  const Region FV4{gx0 - 2, gy0 - 1, TX + 4, TY + 1};
  const Region FU5{gx0 - 1, gy0, TX + 1, TY};
  constexpr int S0 = (TX + 8) * (TY + 8), S1 = (TX + 4) * (TY + 8), S2 = (TX + 4) * (TY + 4),
                S4 = (TX + 4) * TY, ST = TX * TY, SU1 = (TX + 5) * (TY + 8),
                SV2 = (TX + 4) * (TY + 5), SV4 = (TX + 4) * (TY + 1), SU5 = (TX + 1) * TY;
  constexpr int C0 = cdiv(S0, NT), C1 = cdiv(S1, NT), C2 = cdiv(S2, NT), C4 = cdiv(S4, NT),
                CT = cdiv(ST, NT), CU1 = cdiv(SU1, NT), CV2 = cdiv(SV2, NT),
                CV4 = cdiv(SV4, NT), CU5 = cdiv(SU5, NT);

  // ---- LDS
  // cell flags over R0: bit0 mask_uflux, bit1 mask_vflux, bit2 u-update,
  // bit3 v-update, bit4 w column, bit5 inside the array
  __shared__ uint8_t MB[S0];
  __shared__ uint8_t cU1[SU1], cV2[SV2], cV4[SV4], cU5[SU5];
  __shared__ FB fU1[SU1], fV2[SV2], fV4[SV4], fU5[SU5];
  __shared__ double uuL[SU1], vvL[SV2], uuK[SU5], vvK[SV4];
  __shared__ double D1s[S1], D3s[S2], D4s[S4];
  __shared__ TF F0[S0], F1[S1], F3[S2], F4[S4], FL[SU1];
  __shared__ double A1s[ADV ? S1 : 1], A3s[ADV ? S2 : 1], A4s[ADV ? S4 : 1];

  auto inside = [&](int x, int y) { return x >= 0 && x < NX && y >= 0 && y < NY; };
  auto g2 = [&](int x, int y) -> long { return (long)x + (long)NX * y; };
  auto rx = [&](const Region& R, int l) { return R.ox + l % R.w; };
  auto ry = [&](const Region& R, int l) { return R.oy + l / R.w; };
  const TF* fin0 = a.fin + (long)t0 * a.tstride;
  TF* fout0 = a.fout + (long)t0 * a.tstride;

  // ---- prefetch registers (one level ahead)
  TF pf_f[G][C0];
  double pf_ho[C1], pf_uu[CU1], pf_hu[CU1], pf_vv[CV2], pf_hv[CV2];
  double pf_ww[C2], pf_uuK[CU5], pf_huK[CU5], pf_vvK[CV4], pf_hvK[CV4];
  auto prefetch = [&](int L, int K) {
    if (L >= 1 && L <= kmax) {
      const long oL = NXY * (long)L;
#pragma unroll
      for (int q = 0; q < C0; q++) {
        int l = tid + q * NT, x = rx(R0, l), y = ry(R0, l);
        bool ok = l < S0 && inside(x, y);
        long c = oL + g2(x, y);
#pragma unroll
        for (int g = 0; g < G; g++)
          pf_f[g][q] = (ok && g < ntr) ? fin0[(long)g * a.tstride + c] : TF(0);
      }
#pragma unroll
      for (int q = 0; q < C1; q++) {
        int l = tid + q * NT, x = rx(R1, l), y = ry(R1, l);
        pf_ho[q] = (l < S1 && inside(x, y)) ? a.ho[oL + g2(x, y)] : 1.0;
      }
#pragma unroll
      for (int q = 0; q < CU1; q++) {
        int l = tid + q * NT, x = rx(FU1, l), y = ry(FU1, l);
        bool ok = l < SU1 && inside(x, y);
        pf_uu[q] = ok ? a.uu[oL + g2(x, y)] : 0.0;
        pf_hu[q] = ok ? a.hu[oL + g2(x, y)] : 1.0;
      }
#pragma unroll
      for (int q = 0; q < CV2; q++) {
        int l = tid + q * NT, x = rx(FV2, l), y = ry(FV2, l);
        bool ok = l < SV2 && inside(x, y);
        pf_vv[q] = ok ? a.vv[oL + g2(x, y)] : 0.0;
        pf_hv[q] = ok ? a.hv[oL + g2(x, y)] : 1.0;
      }
    }
    if (K >= 1 && K <= kmax) {
      const long oK = NXY * (long)K;
#pragma unroll
      for (int q = 0; q < C2; q++) {
        int l = tid + q * NT, x = rx(R2, l), y = ry(R2, l);
        pf_ww[q] = (l < S2 && inside(x, y)) ? a.ww[oK + g2(x, y)] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < CU5; q++) {
        int l = tid + q * NT, x = rx(FU5, l), y = ry(FU5, l);
        bool ok = l < SU5 && inside(x, y);
        pf_uuK[q] = ok ? a.uu[oK + g2(x, y)] : 0.0;
        pf_huK[q] = ok ? a.hu[oK + g2(x, y)] : 1.0;
      }
#pragma unroll
      for (int q = 0; q < CV4; q++) {
        int l = tid + q * NT, x = rx(FV4, l), y = ry(FV4, l);
        bool ok = l < SV4 && inside(x, y);
        pf_vvK[q] = ok ? a.vv[oK + g2(x, y)] : 0.0;
        pf_hvK[q] = ok ? a.hv[oK + g2(x, y)] : 1.0;
      }
    }
  };

  prefetch(1, -1);
  // cell flags (level independent)
  for (int l = tid; l < S0; l += NT) {
    int x = rx(R0, l), y = ry(R0, l);
    uint8_t b = 0;
    if (inside(x, y)) {
      long c = g2(x, y);
      b |= 32;
      if (a.m_uflux[c]) b |= 1;
      if (a.m_vflux[c]) b |= 2;
      if (a.m_uupd[c] && x >= 1 + isub && x <= NX - 2 - isub) b |= 4;
      if (a.m_vupd[c] && y >= 1 + isub && y <= NY - 2 - isub) b |= 8;
      if (a.az[c] == 1 && x <= NX - 2 && y <= NY - 2) b |= 16;
    }
    MB[l] = b;
  }

  // ---- per-thread state of the owned R2 columns, in registers
  TF f2r[G][C2][4];   // post-v/2 values of levels K-1..K+2 (ring slots 0..3)
  TF wfp[G][C2];      // w flux at the level below the current K
  double d2r[C2][4];  // hi after the first v/2, levels K-1..K+2
  double a2r[C2][4];  // ADV: accumulated adv after the first v/2
  double wwp[C2];     // ww at level K-1
#pragma unroll
  for (int q = 0; q < C2; q++) {
    int l = tid + q * NT, x = rx(R2, l), y = ry(R2, l);
    bool ok = l < S2 && inside(x, y);
    long c = g2(x, y);
    d2r[q][3] = ok ? a.ho[c] : 1.0;
    wwp[q] = ok ? a.ww[c] : 0.0;
    a2r[q][3] = 0.0;
#pragma unroll
    for (int g = 0; g < G; g++) {
      f2r[g][q][3] = (ok && g < ntr) ? fin0[(long)g * a.tstride + c] : TF(0);
      wfp[g][q] = TF(0);
    }
  }
  // level 0 of the outputs
  for (int l = tid; l < ST; l += NT) {
    int x = rx(RT, l), y = ry(RT, l);
    if (!inside(x, y)) continue;
    long c = g2(x, y);
    for (int g = 0; g < ntr; g++) fout0[(long)g * a.tstride + c] = fin0[(long)g * a.tstride + c];
    if (grp == 0 && a.hires) a.hires[c] = a.ho[c];
    if constexpr (ADV) {
      if (a.advres) a.advres[c] = 0.0;
    }
  }

  // hi of the owned elements, kept across the phases of one level
  double d0r[C1], d1r[C1], rd1[C1];
  double d1o[C2], rd2[C2], d3r[C2], rd3[C2], cflw[C2], cflw0[C2], wwK[C2];
  double d3o[C4], d4r[C4], rd4[C4];
  double d4o[CT], d5r[CT], rd5[CT];
  TF pf_cur[G][C0];

  for (int m = 1; m <= kmax + 2; m++) {
    const int L = m, K = m - 2;
    const bool doL = L <= kmax, doK = K >= 1;
    const long oK = NXY * (long)K;

    __syncthreads();  // the previous level is done with every LDS buffer
    // ---- face constants and transports of levels L and K (from prefetch)
    if (doL) {
#pragma unroll
      for (int q = 0; q < CU1; q++) {
        int l = tid + q * NT;
        if (l >= SU1) break;
        int x = rx(FU1, l), y = ry(FU1, l);
        uint8_t mb = MB[R0.idx(x, y)];
        bool act = (mb & 32) && (mb & 1) && x >= isub && x <= NX - 2 - isub;
        bool ll = act && (MB[R0.idx(x - 1, y)] & 1), lr = act && (MB[R0.idx(x + 1, y)] & 1);
        cU1[l] = face_setup<SCH, EXACT>(fU1[l], act, ll, lr, pf_uu[q], pf_hu[q], dth, dx, AH);
        uuL[l] = pf_uu[q];
      }
#pragma unroll
      for (int q = 0; q < CV2; q++) {
        int l = tid + q * NT;
        if (l >= SV2) break;
        int x = rx(FV2, l), y = ry(FV2, l);
        uint8_t mb = MB[R0.idx(x, y)];
        bool act = (mb & 32) && (mb & 2) && y >= isub && y <= NY - 2 - isub;
        bool ll = act && (MB[R0.idx(x, y - 1)] & 2), lr = act && (MB[R0.idx(x, y + 1)] & 2);
        cV2[l] = face_setup<SCH, EXACT>(fV2[l], act, ll, lr, pf_vv[q], pf_hv[q], dth, dy, AH);
        vvL[l] = pf_vv[q];
      }
#pragma unroll
      for (int q = 0; q < C1; q++) d0r[q] = pf_ho[q];
    }
    if (doK) {
#pragma unroll
      for (int q = 0; q < CV4; q++) {
        int l = tid + q * NT;
        if (l >= SV4) break;
        int x = rx(FV4, l), y = ry(FV4, l);
        uint8_t mb = MB[R0.idx(x, y)];
        bool act = (mb & 32) && (mb & 2) && y >= isub && y <= NY - 2 - isub;
        bool ll = act && (MB[R0.idx(x, y - 1)] & 2), lr = act && (MB[R0.idx(x, y + 1)] & 2);
        cV4[l] = face_setup<SCH, EXACT>(fV4[l], act, ll, lr, pf_vvK[q], pf_hvK[q], dth, dy, AH);
        vvK[l] = pf_vvK[q];
      }
#pragma unroll
      for (int q = 0; q < CU5; q++) {
        int l = tid + q * NT;
        if (l >= SU5) break;
        int x = rx(FU5, l), y = ry(FU5, l);
        uint8_t mb = MB[R0.idx(x, y)];
        bool act = (mb & 32) && (mb & 1) && x >= isub && x <= NX - 2 - isub;
        bool ll = act && (MB[R0.idx(x - 1, y)] & 1), lr = act && (MB[R0.idx(x + 1, y)] & 1);
        cU5[l] = face_setup<SCH, EXACT>(fU5[l], act, ll, lr, pf_uuK[q], pf_huK[q], dth, dx, AH);
        uuK[l] = pf_uuK[q];
      }
#pragma unroll
      for (int q = 0; q < C2; q++) wwK[q] = pf_ww[q];
    }
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
      for (int q = 0; q < C0; q++) pf_cur[g][q] = pf_f[g][q];
    prefetch(L + 1, K + 1);
    // ring shift of the owned columns: slot 3 receives level L below
#pragma unroll
    for (int q = 0; q < C2; q++) {
      d2r[q][0] = d2r[q][1]; d2r[q][1] = d2r[q][2]; d2r[q][2] = d2r[q][3];
      if constexpr (ADV) {
        a2r[q][0] = a2r[q][1]; a2r[q][1] = a2r[q][2]; a2r[q][2] = a2r[q][3];
      }
#pragma unroll
      for (int g = 0; g < G; g++) {
        f2r[g][q][0] = f2r[g][q][1]; f2r[g][q][1] = f2r[g][q][2]; f2r[g][q][2] = f2r[g][q][3];
      }
    }
    __syncthreads();

    // ---- shared hi: u/2 at level L (R1)
    if (doL) {
#pragma unroll
      for (int q = 0; q < C1; q++) {
        int l = tid + q * NT;
        if (l >= S1) break;
        int x = rx(R1, l), y = ry(R1, l);
        double D0 = d0r[q], D1 = D0;
        if (MB[R0.idx(x, y)] & 4)
          D1 = D0 - dth * (uuL[FU1.idx(x, y)] * dy - uuL[FU1.idx(x - 1, y)] * dy) * ard1;
        d1r[q] = D1;
        if constexpr (!EXACT) rd1[q] = 1.0 / D1;
        D1s[l] = D1;
      }
    }
    __syncthreads();
    // ---- shared hi: v/2 at level L and w at level K (owned R2 columns)
#pragma unroll
    for (int q = 0; q < C2; q++) {
      int l = tid + q * NT;
      if (l >= S2) break;
      int x = rx(R2, l), y = ry(R2, l);
      uint8_t mb = MB[R0.idx(x, y)];
      if (doL) {
        double D1 = D1s[R1.idx(x, y)], D2 = D1;
        if (mb & 8) D2 = D1 - dth * (vvL[FV2.idx(x, y)] * dx - vvL[FV2.idx(x, y - 1)] * dx) * ard1;
        d1o[q] = D1;
        d2r[q][3] = D2;
        if constexpr (!EXACT) rd2[q] = 1.0 / D2;
      }
      if (doK) {
        // ring: slot 0 = level K-1, slot 1 = K, slot 2 = K+1
        double D2 = d2r[q][1], D3 = D2;
        if ((mb & 16) && K <= kmax - kshift) D3 = D2 - dt * (wwK[q] - wwp[q]);
        d3r[q] = D3;
        if constexpr (!EXACT) rd3[q] = 1.0 / D3;
        D3s[l] = D3;
        cflw[q] = fabs(wwK[q]) * dt / (0.5 * (D2 + d2r[q][2]));
        cflw0[q] = fabs(wwp[q]) * dt / (0.5 * (d2r[q][0] + D2));
      }
    }
    if (doK) {
      __syncthreads();
#pragma unroll
      for (int q = 0; q < C4; q++) {
        int l = tid + q * NT;
        if (l >= S4) break;
        int x = rx(R4, l), y = ry(R4, l);
        double D3 = D3s[R2.idx(x, y)], D4 = D3;
        if (MB[R0.idx(x, y)] & 8)
          D4 = D3 - dth * (vvK[FV4.idx(x, y)] * dx - vvK[FV4.idx(x, y - 1)] * dx) * ard1;
        d3o[q] = D3;
        d4r[q] = D4;
        if constexpr (!EXACT) rd4[q] = 1.0 / D4;
        D4s[l] = D4;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < CT; q++) {
        int l = tid + q * NT;
        if (l >= ST) break;
        int x = rx(RT, l), y = ry(RT, l);
        double D4 = D4s[R4.idx(x, y)], D5 = D4;
        if (MB[R0.idx(x, y)] & 4)
          D5 = D4 - dth * (uuK[FU5.idx(x, y)] * dy - uuK[FU5.idx(x - 1, y)] * dy) * ard1;
        d4o[q] = D4;
        d5r[q] = D5;
        if constexpr (!EXACT) rd5[q] = 1.0 / D5;
        if (grp == 0 && a.hires && inside(x, y)) a.hires[oK + g2(x, y)] = D5;
      }
    }

#pragma unroll
    for (int g = 0; g < G; g++) {
      if (g >= ntr) break;
      TF* fout = fout0 + (long)g * a.tstride;
      if (doL) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < C0; q++) {
          int l = tid + q * NT;
          if (l < S0) F0[l] = pf_cur[g][q];
        }
        __syncthreads();
        // u/2 fluxes on faces FU1 (adv_split_u.F90:219-265)
#pragma unroll
        for (int q = 0; q < CU1; q++) {
          int l = tid + q * NT;
          if (l >= SU1) break;
          int x = rx(FU1, l), y = ry(FU1, l);
          FL[l] = face_flux<TF, SCH, EXACT>(cU1[l], fU1[l], F0, R0.idx(x, y), 1, AH, dx);
        }
        __syncthreads();
        // u/2 update on R1 (adv_split_u.F90:267-286)
#pragma unroll
        for (int q = 0; q < C1; q++) {
          int l = tid + q * NT;
          if (l >= S1) break;
          int x = rx(R1, l), y = ry(R1, l);
          int l0 = R0.idx(x, y);
          TF f0 = F0[l0], f1 = f0;
          double advn = 0.0;
          if (MB[l0] & 4)
            f1 = cell_update<TF, EXACT>(d0r[q], d1r[q], rd1[q], f0, FL[FU1.idx(x, y)],
                                        FL[FU1.idx(x - 1, y)], cdu, dy, ard1, dt, &advn);
          F1[l] = f1;
          if constexpr (ADV) A1s[l] = 0.0 + advn;
        }
        __syncthreads();
        // v/2 fluxes on faces FV2 (adv_split_v.F90:77-119)
#pragma unroll
        for (int q = 0; q < CV2; q++) {
          int l = tid + q * NT;
          if (l >= SV2) break;
          int x = rx(FV2, l), y = ry(FV2, l);
          FL[l] = face_flux<TF, SCH, EXACT>(cV2[l], fV2[l], F1, R1.idx(x, y), R1.w, AH, dy);
        }
        __syncthreads();
      }
      // v/2 update at level L and w at level K on the owned R2 columns
#pragma unroll
      for (int q = 0; q < C2; q++) {
        int l = tid + q * NT;
        if (l >= S2) break;
        int x = rx(R2, l), y = ry(R2, l);
        uint8_t mb = MB[R0.idx(x, y)];
        if (doL) {
          int l1 = R1.idx(x, y);
          TF f1 = F1[l1], f2 = f1;
          double advn = 0.0;
          if (mb & 8)
            f2 = cell_update<TF, EXACT>(d1o[q], d2r[q][3], rd2[q], f1, FL[FV2.idx(x, y)],
                                        FL[FV2.idx(x, y - 1)], cdv, dx, ard1, dt, &advn);
          f2r[g][q][3] = f2;
          if constexpr (ADV) a2r[q][3] = A1s[l1] + advn;
        }
        if (doK) {
          TF f2 = f2r[g][q][1], f3 = f2;
          double advn = 0.0;
          if (mb & 16) {
            TF wprev = wfp[g][q];
            if (K == 1)
              wprev = kshift ? wflux<TF, SCH, EXACT>(0, kshift, kmax, wwp[q], TF(0),
                                                     f2r[g][q][0], f2, f2r[g][q][2], cflw0[q])
                             : TF(0);
            TF wk = TF(0);
            if (K <= kmax - 1)
              wk = wflux<TF, SCH, EXACT>(K, kshift, kmax, wwK[q], f2r[g][q][0], f2,
                                         f2r[g][q][2], f2r[g][q][3], cflw[q]);
            wfp[g][q] = wk;
            if (K <= kmax - kshift) {
              if constexpr (EXACT) {
                advn = 1.0 * (wk - wprev);
                f3 = (d2r[q][1] * f2 - dt * advn) / d3r[q];
              } else {
                f3 = (TF)((d2r[q][1] * (double)f2 - dt * (double)(wk - wprev)) * rd3[q]);
              }
            }
          }
          F3[l] = f3;
          if constexpr (ADV) A3s[l] = a2r[q][1] + advn;
        }
      }
      if (doK) {
        __syncthreads();
        // second v/2: fluxes on FV4 from F3, update on R4
#pragma unroll
        for (int q = 0; q < CV4; q++) {
          int l = tid + q * NT;
          if (l >= SV4) break;
          int x = rx(FV4, l), y = ry(FV4, l);
          FL[l] = face_flux<TF, SCH, EXACT>(cV4[l], fV4[l], F3, R2.idx(x, y), R2.w, AH, dy);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < C4; q++) {
          int l = tid + q * NT;
          if (l >= S4) break;
          int x = rx(R4, l), y = ry(R4, l);
          int l2 = R2.idx(x, y);
          TF f3 = F3[l2], f4 = f3;
          double advn = 0.0;
          if (MB[R0.idx(x, y)] & 8)
            f4 = cell_update<TF, EXACT>(d3o[q], d4r[q], rd4[q], f3, FL[FV4.idx(x, y)],
                                        FL[FV4.idx(x, y - 1)], cdv, dx, ard1, dt, &advn);
          F4[l] = f4;
          if constexpr (ADV) A4s[l] = A3s[l2] + advn;
        }
        __syncthreads();
        // second u/2: fluxes on FU5 from F4, update on the tile, store
#pragma unroll
        for (int q = 0; q < CU5; q++) {
          int l = tid + q * NT;
          if (l >= SU5) break;
          int x = rx(FU5, l), y = ry(FU5, l);
          FL[l] = face_flux<TF, SCH, EXACT>(cU5[l], fU5[l], F4, R4.idx(x, y), 1, AH, dx);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < CT; q++) {
          int l = tid + q * NT;
          if (l >= ST) break;
          int x = rx(RT, l), y = ry(RT, l);
          if (!inside(x, y)) continue;
          int l4 = R4.idx(x, y);
          TF f4 = F4[l4], f5 = f4;
          double advn = 0.0;
          if (MB[R0.idx(x, y)] & 4)
            f5 = cell_update<TF, EXACT>(d4o[q], d5r[q], rd5[q], f4, FL[FU5.idx(x, y)],
                                        FL[FU5.idx(x - 1, y)], cdu, dy, ard1, dt, &advn);
          fout[oK + g2(x, y)] = f5;
          if constexpr (ADV) {
            if (a.advres) a.advres[oK + g2(x, y)] = A4s[l4] + advn;
          }
        }
      }
    }
    if (doK) {
#pragma unroll
      for (int q = 0; q < C2; q++) wwp[q] = wwK[q];
    }
  }
}

template <typename TF, int SCH, int G, bool EXACT, bool ADV, int TX, int TY, int NT>
hipError_t launch_one(hipStream_t st, FusedArgs<TF> A) {
  A.ngroups = (A.ntracers + G - 1) / G;
  long nb = A.ntiles * A.ngroups;
  long grid = ((nb + 7) / 8) * 8;
  hipLaunchKernelGGL((k_fused_halfsplit<TF, SCH, G, EXACT, ADV, TX, TY, NT>),
                     dim3((unsigned)grid), dim3(NT), 0, st, A);
  return hipGetLastError();
}

// one scheme at compile time (hscheme == vscheme)
template <typename TF, int G, bool EXACT, bool ADV, int TX, int TY, int NT>
hipError_t dispatch_scheme(hipStream_t st, const FusedArgs<TF>& a) {
  if (a.hscheme != a.vscheme) return hipErrorInvalidValue;
  switch (a.hscheme) {
    case P2_PDM: return launch_one<TF, P2_PDM, G, EXACT, ADV, TX, TY, NT>(st, a);
    case SUPERBEE: return launch_one<TF, SUPERBEE, G, EXACT, ADV, TX, TY, NT>(st, a);
    case MUSCL: return launch_one<TF, MUSCL, G, EXACT, ADV, TX, TY, NT>(st, a);
    case P2: return launch_one<TF, P2, G, EXACT, ADV, TX, TY, NT>(st, a);
    case UPSTREAM: return launch_one<TF, UPSTREAM, G, EXACT, ADV, TX, TY, NT>(st, a);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gt
