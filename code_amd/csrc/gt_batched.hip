// gt_batched.hip -- fused HALFSPLIT transport step: u/2 + v/2 + w + v/2 + u/2
// (src/3d/advection_3d.F90:439-547 of getm-model/code) in ONE pass over HBM.
//
// Design (MI355X): each workgroup owns a TX x TY tile of water columns and
// marches up the column.  At march iteration m it
//   * runs the two leading horizontal sub-steps (u/2, v/2) on level L = m,
//     staging f in LDS on the tile plus a 4-cell ghost ring (the stencil
//     radius of two consecutive 2-cell sub-steps);
//   * keeps the post-v/2 values of the last four levels in an LDS ring, which
//     is exactly the vertical stencil of the w sub-step (k-1..k+2);
//   * runs w, then the trailing v/2 and u/2, on level K = m-2 and writes the
//     result.
// The layer-height trajectory hi (FP64) is shared by every tracer of the
// block and computed once per level; each tracer adds only its own f traffic.
// So the compulsory traffic of one step is: f read + written per tracer, and
// uu, vv, ww, hu, hv, ho read once (hn is not needed by the update).
//
// Two arithmetic flavours:
//   EXACT (TF=double): the reference's expressions and order, so the result
//     is bit-identical to the per-sub-step kernels and the CPU restatement.
//   mixed (TF=float): tracer and flux arithmetic in FP32; hi and the cell
//     update (Dio*f - dt*advn)/Di in FP64, rounded to FP32 once per sub-step.
//
// Which cells each sub-step updates follows the reference exactly (the isub /
// jsub bounds of adv_split_u/v.F90 and the column range of adv_split_w.F90),
// so cells the reference leaves alone pass through unchanged.
#include "gt_batched.h"

namespace gt {

struct Region {
  int ox, oy, w, h;
  __device__ __forceinline__ int idx(int x, int y) const { return (x - ox) + w * (y - oy); }
  __device__ __forceinline__ int size() const { return w * h; }
};

template <typename TF>
struct Arith;

// reference flux in x at face (gx, gy) of level-slice `f` held in region R.
// mask arrays m* are held in LDS over region RM.  Returns the uflux of
// adv_split_u.F90:219-265 (0 where mask_flux is false).
template <typename TF, bool EXACT>
__device__ __forceinline__ TF face_flux(int scheme, const TF* fL, const Region& R, int x, int y,
                                        int dxs, int dys, const uint8_t* mflux, const Region& RM,
                                        double vel, double Dvel, double dti, double dlen,
                                        double AH) {
  // (dxs,dys) is the unit step along the sweep direction
  if (!mflux[RM.idx(x, y)]) return TF(0);
  TF fu, fuu = TF(0), fd = TF(0);
  bool use_limiter = false;
  double cfl = 0.0;
  if (vel > 0.0) {
    fu = fL[R.idx(x, y)];
    if (scheme != UPSTREAM) use_limiter = mflux[RM.idx(x - dxs, y - dys)];
    if (use_limiter) {
      cfl = vel / Dvel * dti / dlen;
      fuu = fL[R.idx(x - dxs, y - dys)];
      fd = fL[R.idx(x + dxs, y + dys)];
    }
  } else {
    fu = fL[R.idx(x + dxs, y + dys)];
    if (scheme != UPSTREAM) use_limiter = mflux[RM.idx(x + dxs, y + dys)];
    if (use_limiter) {
      cfl = -vel / Dvel * dti / dlen;
      fuu = fL[R.idx(x + 2 * dxs, y + 2 * dys)];
      fd = fL[R.idx(x, y)];
    }
  }
  if (use_limiter) fu = reconstruct<TF>(scheme, (TF)cfl, fuu, fu, fd);
  TF fl = TF(vel) * fu;
  if (AH > 0.0) {
    if constexpr (EXACT) {
      fl = fl - AH * Dvel * (fL[R.idx(x + dxs, y + dys)] - fL[R.idx(x, y)]) / dlen;
    } else {
      fl = fl - TF(AH * Dvel / dlen) * (fL[R.idx(x + dxs, y + dys)] - fL[R.idx(x, y)]);
    }
  }
  return fl;
}

// one cell of a horizontal sub-step: the new tracer value
template <typename TF, bool EXACT>
__device__ __forceinline__ TF cell_update(double Dio, double Dn, TF fo, TF fl_c, TF fl_m,
                                          double splitfac, double dlen, double ard1, double dt,
                                          double* advn_out) {
  if constexpr (EXACT) {
    double advn = splitfac * (fl_c * dlen - fl_m * dlen) * ard1;
    *advn_out = advn;
    return (Dio * fo - dt * advn) / Dn;
  } else {
    TF advn = TF(splitfac * dlen * ard1) * (fl_c - fl_m);
    *advn_out = advn;
    return (TF)((Dio * (double)fo - dt * (double)advn) / Dn);
  }
}

// vertical flux at level k of a column (adv_split_w.F90:249-277); the window
// holds f at k-1, k, k+1, k+2
template <typename TF>
__device__ __forceinline__ TF wflux_at(int scheme, int k, int kshift, int kmax, double w,
                                       TF fkm1, TF fk, TF fk1, TF fk2, double hk, double hk1,
                                       double dti) {
  TF fu, fuu = TF(0), fd = TF(0);
  bool use_limiter = false;
  if (w > 0.0) {
    fu = fk;
    if (scheme != UPSTREAM) use_limiter = k > 1 - kshift;
    if (use_limiter) {
      fuu = fkm1;
      fd = fk1;
    }
  } else {
    fu = fk1;
    if (scheme != UPSTREAM) use_limiter = k < kmax - 1;
    if (use_limiter) {
      fuu = fk2;
      fd = fk;
    }
  }
  if (use_limiter) {
    double cfl = fabs(w) * dti / (0.5 * (hk + hk1));
    fu = reconstruct<TF>(scheme, (TF)(cfl * 1.0), fuu, fu, fd);
  }
  return TF(w) * fu;
}

template <typename TF, int G, bool EXACT, bool ADV, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_fused_halfsplit(FusedArgs<TF> A) {
  // ---- block -> (domain, tile, tracer group), XCD-aware: blocks
  // b, b+8, b+16, ... run on one XCD, so they get consecutive work items and
  // neighbouring tiles (which share ghost cells) meet in that XCD's L2
  const long nb = A.ntiles * A.ngroups;
  const long per = (nb + 7) / 8;
  const long q = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (q >= nb) return;
  const long gtile = q / A.ngroups;
  const int grp = (int)(q % A.ngroups);
  int di = 0;
  while (di + 1 < A.ndom && A.doms[di + 1].tile_begin <= gtile) di++;
  const FusedDomain<TF>& a = A.doms[di];
  const int tile = (int)(gtile - a.tile_begin);
  const int gx0 = (tile % a.ntx) * TX, gy0 = (tile / a.ntx) * TY;
  const int NX = a.d.NX, NY = a.d.NY, kmax = a.d.kmax;
  const long NXY = a.d.NXY;
  const int tid = threadIdx.x;
  const int scheme = A.hscheme, vscheme = A.vscheme, kshift = A.kshift;
  const int isub = scheme == UPSTREAM ? 0 : 1;
  const double dt = A.dt, dth = 0.5 * A.dt, dx = a.dx, dy = a.dy, ard1 = a.ard1, AH = A.AH;
  const int t0 = grp * G;
  const int ntr = min(G, A.ntracers - t0);

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
                S4 = (TX + 4) * TY, SFL = (TX + 5) * (TY + 8);

  // ---- LDS
  __shared__ uint8_t m_uf[S0], m_vf[S0], m_uu[S0], m_vu[S0], m_az[S0];
  __shared__ double D0s[S1], D1s[S1], D2r[4][S2], D3s[S2], D4s[S4];
  __shared__ TF F0[S0], F1[S1], F2r[4][G][S2], WF[G][S2], F3[S2], F4[S4], FL[SFL];
  __shared__ double A1s[ADV ? S1 : 1], A2r[ADV ? 3 : 1][ADV ? S2 : 1], A3s[ADV ? S2 : 1],
      A4s[ADV ? S4 : 1];

  auto inside = [&](int x, int y) { return x >= 0 && x < NX && y >= 0 && y < NY; };
  auto g2 = [&](int x, int y) -> long { return (long)x + (long)NX * y; };

  // masks over R0
  for (int l = tid; l < S0; l += NT) {
    int x = R0.ox + l % R0.w, y = R0.oy + l / R0.w;
    bool in = inside(x, y);
    long c = g2(x, y);
    m_uf[l] = in ? a.m_uflux[c] : 0;
    m_vf[l] = in ? a.m_vflux[c] : 0;
    m_uu[l] = in ? a.m_uupd[c] : 0;
    m_vu[l] = in ? a.m_vupd[c] : 0;
    m_az[l] = in ? (a.az[c] == 1) : 0;
  }
  // level 0: never advanced horizontally; seeds the ring for the w stencil
  for (int l = tid; l < S2; l += NT) {
    int x = R2.ox + l % R2.w, y = R2.oy + l / R2.w;
    bool in = inside(x, y);
    long c = g2(x, y);
    D2r[0][l] = in ? a.ho[c] : 1.0;
    if constexpr (ADV) A2r[0][l] = 0.0;
    for (int g = 0; g < ntr; g++) F2r[0][g][l] = in ? a.fin[(long)(t0 + g) * a.tstride + c] : TF(0);
  }
  // level 0 of the outputs
  for (int l = tid; l < TX * TY; l += NT) {
    int x = RT.ox + l % TX, y = RT.oy + l / TX;
    if (!inside(x, y)) continue;
    long c = g2(x, y);
    for (int g = 0; g < ntr; g++)
      a.fout[(long)(t0 + g) * a.tstride + c] = a.fin[(long)(t0 + g) * a.tstride + c];
    if (grp == 0 && a.hires) a.hires[c] = a.ho[c];
    if constexpr (ADV) {
      if (a.advres) a.advres[c] = 0.0;
    }
  }
  __syncthreads();

  for (int m = 1; m <= kmax + 2; m++) {
    const int L = m, K = m - 2;
    const bool doL = L <= kmax, doK = K >= 1;
    const long oL = NXY * (long)L, oK = NXY * (long)K;
    const int sL = L % 4, sK = (K + 4) % 4, sK1 = (K + 1) % 4;
    const int aL = L % 3, aK = (K + 3) % 3;

    // ---- shared: hi through u/2 and v/2 at level L
    if (doL) {
      for (int l = tid; l < S1; l += NT) {
        int x = R1.ox + l % R1.w, y = R1.oy + l / R1.w;
        double D0 = 1.0, D1 = 1.0;
        if (inside(x, y)) {
          long c = oL + g2(x, y);
          D0 = a.ho[c];
          D1 = D0;
          if (x >= 1 + isub && x <= NX - 2 - isub && m_uu[R0.idx(x, y)])
            D1 = D0 - dth * (a.uu[c] * dy - a.uu[c - 1] * dy) * ard1;
        }
        D0s[l] = D0;
        D1s[l] = D1;
      }
      __syncthreads();
      for (int l = tid; l < S2; l += NT) {
        int x = R2.ox + l % R2.w, y = R2.oy + l / R2.w;
        double D1 = D1s[R1.idx(x, y)], D2 = D1;
        if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)]) {
          long c = oL + g2(x, y);
          D2 = D1 - dth * (a.vv[c] * dx - a.vv[c - NX] * dx) * ard1;
        }
        D2r[sL][l] = D2;
      }
    }
    // ---- shared: hi through w, v/2 at level K (u/2 is done per tracer)
    if (doK) {
      __syncthreads();
      for (int l = tid; l < S2; l += NT) {
        int x = R2.ox + l % R2.w, y = R2.oy + l / R2.w;
        double D2 = D2r[sK][l], D3 = D2;
        if (inside(x, y) && x <= NX - 2 && y <= NY - 2 && m_az[R0.idx(x, y)] &&
            K <= kmax - kshift && vscheme != NOADV) {
          long c = oK + g2(x, y);
          D3 = D2 - dt * (a.ww[c] - a.ww[c - NXY]);
        }
        D3s[l] = D3;
      }
      __syncthreads();
      for (int l = tid; l < S4; l += NT) {
        int x = R4.ox + l % R4.w, y = R4.oy + l / R4.w;
        double D3 = D3s[R2.idx(x, y)], D4 = D3;
        if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)]) {
          long c = oK + g2(x, y);
          D4 = D3 - dth * (a.vv[c] * dx - a.vv[c - NX] * dx) * ard1;
        }
        D4s[l] = D4;
      }
    }
    __syncthreads();

    for (int g = 0; g < ntr; g++) {
      const TF* fin = a.fin + (long)(t0 + g) * a.tstride;
      TF* fout = a.fout + (long)(t0 + g) * a.tstride;
      // ================= level L: u/2 then v/2 =================
      if (doL) {
        for (int l = tid; l < S0; l += NT) {
          int x = R0.ox + l % R0.w, y = R0.oy + l / R0.w;
          F0[l] = inside(x, y) ? fin[oL + g2(x, y)] : TF(0);
        }
        __syncthreads();
        // u fluxes on faces FU1 (adv_split_u.F90:219-265)
        for (int l = tid; l < SFL; l += NT) {
          int x = FU1.ox + l % FU1.w, y = FU1.oy + l / FU1.w;
          TF fl = TF(0);
          if (inside(x, y) && x >= isub && x <= NX - 2 - isub) {
            long c = oL + g2(x, y);
            fl = face_flux<TF, EXACT>(scheme, F0, R0, x, y, 1, 0, m_uf, R0, a.uu[c], a.hu[c],
                                      dth, dx, AH);
          }
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < S1; l += NT) {
          int x = R1.ox + l % R1.w, y = R1.oy + l / R1.w;
          TF f0 = F0[R0.idx(x, y)], f1 = f0;
          double advn = 0.0;
          if (inside(x, y) && x >= 1 + isub && x <= NX - 2 - isub && m_uu[R0.idx(x, y)]) {
            f1 = cell_update<TF, EXACT>(D0s[l], D1s[l], f0, FL[FU1.idx(x, y)],
                                        FL[FU1.idx(x - 1, y)], 0.5, dy, ard1, dt, &advn);
          }
          F1[l] = f1;
          if constexpr (ADV) A1s[l] = 0.0 + advn;
        }
        __syncthreads();
        // v fluxes on faces FV2 (adv_split_v.F90:77-119)
        for (int l = tid; l < (TX + 4) * (TY + 5); l += NT) {
          int x = FV2.ox + l % FV2.w, y = FV2.oy + l / FV2.w;
          TF fl = TF(0);
          if (inside(x, y) && y >= isub && y <= NY - 2 - isub) {
            long c = oL + g2(x, y);
            fl = face_flux<TF, EXACT>(scheme, F1, R1, x, y, 0, 1, m_vf, R0, a.vv[c], a.hv[c],
                                      dth, dy, AH);
          }
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < S2; l += NT) {
          int x = R2.ox + l % R2.w, y = R2.oy + l / R2.w;
          int l1 = R1.idx(x, y);
          TF f1 = F1[l1], f2 = f1;
          double advn = 0.0;
          if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)]) {
            f2 = cell_update<TF, EXACT>(D1s[l1], D2r[sL][l], f1, FL[FV2.idx(x, y)],
                                        FL[FV2.idx(x, y - 1)], 0.5, dx, ard1, dt, &advn);
          }
          F2r[L % 4][g][l] = f2;
          if constexpr (ADV) A2r[aL][l] = A1s[l1] + advn;
        }
      }
      // ================= level K: w, then v/2 and u/2 =================
      if (doK) {
        __syncthreads();
        for (int l = tid; l < S2; l += NT) {
          int x = R2.ox + l % R2.w, y = R2.oy + l / R2.w;
          TF f2 = F2r[K % 4][g][l], f3 = f2;
          double advn = 0.0;
          if (inside(x, y) && x <= NX - 2 && y <= NY - 2 && m_az[R0.idx(x, y)] &&
              vscheme != NOADV) {
            long c = g2(x, y);
            const double dti = dt;  // splitfac 1 for w
            TF wprev;
            if (K == 1) {
              if (kshift) {
                wprev = wflux_at<TF>(vscheme, 0, kshift, kmax, a.ww[c], TF(0), F2r[0][g][l],
                                     F2r[1][g][l], F2r[2][g][l], D2r[0][l], D2r[1][l], dti);
              } else {
                wprev = TF(0);
              }
            } else {
              wprev = WF[g][l];
            }
            TF wk = TF(0);
            if (K <= kmax - 1) {
              TF fkm1 = F2r[(K + 3) % 4][g][l];
              TF fk1 = F2r[(K + 1) % 4][g][l];
              TF fk2 = (K + 2 <= kmax) ? F2r[(K + 2) % 4][g][l] : TF(0);
              wk = wflux_at<TF>(vscheme, K, kshift, kmax, a.ww[oK + c], fkm1, f2, fk1, fk2,
                                D2r[sK][l], D2r[sK1][l], dti);
            }
            WF[g][l] = wk;
            if (K <= kmax - kshift) {
              double Dio = D2r[sK][l], Dn = D3s[l];
              if constexpr (EXACT) {
                advn = 1.0 * (wk - wprev);
                f3 = (Dio * f2 - dt * advn) / Dn;
              } else {
                TF an = wk - wprev;
                advn = an;
                f3 = (TF)((Dio * (double)f2 - dt * (double)an) / Dn);
              }
            }
          }
          F3[l] = f3;
          if constexpr (ADV) A3s[l] = A2r[aK][l] + advn;
        }
        __syncthreads();
        // v/2 fluxes on FV4 from F3
        for (int l = tid; l < (TX + 4) * (TY + 1); l += NT) {
          int x = FV4.ox + l % FV4.w, y = FV4.oy + l / FV4.w;
          TF fl = TF(0);
          if (inside(x, y) && y >= isub && y <= NY - 2 - isub) {
            long c = oK + g2(x, y);
            fl = face_flux<TF, EXACT>(scheme, F3, R2, x, y, 0, 1, m_vf, R0, a.vv[c], a.hv[c],
                                      dth, dy, AH);
          }
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < S4; l += NT) {
          int x = R4.ox + l % R4.w, y = R4.oy + l / R4.w;
          int l2 = R2.idx(x, y);
          TF f3 = F3[l2], f4 = f3;
          double advn = 0.0;
          if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)]) {
            f4 = cell_update<TF, EXACT>(D3s[l2], D4s[l], f3, FL[FV4.idx(x, y)],
                                        FL[FV4.idx(x, y - 1)], 0.5, dx, ard1, dt, &advn);
          }
          F4[l] = f4;
          if constexpr (ADV) A4s[l] = A3s[l2] + advn;
        }
        __syncthreads();
        // u/2 fluxes on FU5 from F4
        for (int l = tid; l < (TX + 1) * TY; l += NT) {
          int x = FU5.ox + l % FU5.w, y = FU5.oy + l / FU5.w;
          TF fl = TF(0);
          if (inside(x, y) && x >= isub && x <= NX - 2 - isub) {
            long c = oK + g2(x, y);
            fl = face_flux<TF, EXACT>(scheme, F4, R4, x, y, 1, 0, m_uf, R0, a.uu[c], a.hu[c],
                                      dth, dx, AH);
          }
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < TX * TY; l += NT) {
          int x = RT.ox + l % TX, y = RT.oy + l / TX;
          if (!inside(x, y)) continue;
          int l4 = R4.idx(x, y);
          TF f4 = F4[l4], f5 = f4;
          double D4 = D4s[l4], D5 = D4;
          double advn = 0.0;
          long c = oK + g2(x, y);
          if (x >= 1 + isub && x <= NX - 2 - isub && m_uu[R0.idx(x, y)]) {
            D5 = D4 - dth * (a.uu[c] * dy - a.uu[c - 1] * dy) * ard1;
            f5 = cell_update<TF, EXACT>(D4, D5, f4, FL[FU5.idx(x, y)], FL[FU5.idx(x - 1, y)],
                                        0.5, dy, ard1, dt, &advn);
          }
          fout[c] = f5;
          if (g == 0 && grp == 0 && a.hires) a.hires[c] = D5;
          if constexpr (ADV) {
            if (a.advres) a.advres[c] = A4s[l4] + advn;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
template <typename TF, int G, bool EXACT, bool ADV, int TX, int TY, int NT>
static hipError_t launch_one(hipStream_t st, FusedArgs<TF> A) {
  A.ngroups = (A.ntracers + G - 1) / G;
  long nb = A.ntiles * A.ngroups;
  long grid = ((nb + 7) / 8) * 8;
  hipLaunchKernelGGL((k_fused_halfsplit<TF, G, EXACT, ADV, TX, TY, NT>), dim3((unsigned)grid),
                     dim3(NT), 0, st, A);
  return hipGetLastError();
}

void fused_tile_shape(int tile, int* TX, int* TY) {
  switch (tile) {
    case 1: *TX = 64; *TY = 4; break;
    case 2: *TX = 16; *TY = 8; break;
    default: *TX = 32; *TY = 8; break;
  }
}

hipError_t launch_fused_f64(hipStream_t st, const FusedArgs<double>& a, int tile, bool adv) {
  switch (tile) {
    case 1:
      return adv ? launch_one<double, 1, true, true, 64, 4, 256>(st, a)
                 : launch_one<double, 1, true, false, 64, 4, 256>(st, a);
    case 2:
      return adv ? launch_one<double, 1, true, true, 16, 8, 256>(st, a)
                 : launch_one<double, 1, true, false, 16, 8, 256>(st, a);
    default:
      return adv ? launch_one<double, 1, true, true, 32, 8, 256>(st, a)
                 : launch_one<double, 1, true, false, 32, 8, 256>(st, a);
  }
}

hipError_t launch_fused_f32(hipStream_t st, const FusedArgs<float>& a, int tile, int group) {
#define GT_F32(G_, TX_, TY_) launch_one<float, G_, false, false, TX_, TY_, 256>(st, a)
  switch (group) {
    case 1:
      return tile == 1 ? GT_F32(1, 64, 4) : tile == 2 ? GT_F32(1, 16, 8) : GT_F32(1, 32, 8);
    case 2:
      return tile == 1 ? GT_F32(2, 64, 4) : tile == 2 ? GT_F32(2, 16, 8) : GT_F32(2, 32, 8);
    case 6:
      return tile == 1 ? GT_F32(6, 64, 4) : tile == 2 ? GT_F32(6, 16, 8) : GT_F32(6, 32, 8);
    default:
      return tile == 1 ? GT_F32(3, 64, 4) : tile == 2 ? GT_F32(3, 16, 8) : GT_F32(3, 32, 8);
  }
#undef GT_F32
}

}  // namespace gt
