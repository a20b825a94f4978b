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

// reference flux at face (x, y) of the level slice fL (held over region R);
// (dxs, dys) is the unit step along the sweep.  The masks live in LDS over
// RM; vel/Dvel are the face's transport and layer height.  This is the uflux
// (vflux) of adv_split_u.F90:219-265 (adv_split_v.F90:77-119), 0 where
// mask_flux is false.
template <typename TF, bool EXACT>
__device__ __forceinline__ TF face_flux(int scheme, const TF* fL, const Region& R, int x, int y,
                                        int dxs, int dys, const uint8_t* mflux, const Region& RM,
                                        double vel, double Dvel, double dti, double dlen,
                                        double AH) {
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

constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// The march is software-pipelined: at the top of iteration m the data of
// levels L = m and K = m-2 that was prefetched into registers during the
// previous iteration is written to LDS, and the loads for the next iteration
// are issued before any compute, so HBM latency overlaps the compute of a
// whole level.  Every LDS loop maps element l to thread l mod NT, the same
// mapping in every phase, which lets a thread reuse what it wrote itself
// without a barrier (the post-v/2 ring feeding the w sub-step).
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
                S4 = (TX + 4) * TY, ST = TX * TY, SU1 = (TX + 5) * (TY + 8),
                SV2 = (TX + 4) * (TY + 5), SV4 = (TX + 4) * (TY + 1), SU5 = (TX + 1) * TY;
  constexpr int C0 = cdiv(S0, NT), C1 = cdiv(S1, NT), C2 = cdiv(S2, NT), CU1 = cdiv(SU1, NT),
                CV2 = cdiv(SV2, NT), CV4 = cdiv(SV4, NT), CU5 = cdiv(SU5, NT);

  // ---- LDS
  __shared__ uint8_t m_uf[S0], m_vf[S0], m_uu[S0], m_vu[S0], m_az[S0];
  __shared__ double uuL[SU1], huL[SU1], vvL[SV2], hvL[SV2], hoL[S1];
  __shared__ double uuK[SU5], huK[SU5], vvK[SV4], hvK[SV4], wwK[2][S2];
  __shared__ double D1s[S1], D2r[4][S2], D3s[S2], D4s[S4];
  __shared__ TF F0[G][S0], F1[S1], F2r[4][G][S2], WF[G][S2], F3[S2], F4[S4], FL[SU1];
  __shared__ double A1s[ADV ? S1 : 1], A2r[ADV ? 3 : 1][ADV ? S2 : 1], A3s[ADV ? S2 : 1],
      A4s[ADV ? S4 : 1];

  auto inside = [&](int x, int y) { return x >= 0 && x < NX && y >= 0 && y < NY; };
  auto g2 = [&](int x, int y) -> long { return (long)x + (long)NX * y; };
  auto rxy = [&](const Region& R, int l, int& x, int& y) {
    x = R.ox + l % R.w;
    y = R.oy + l / R.w;
  };
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
        int l = tid + q * NT, x, y;
        rxy(R0, l, x, y);
        bool ok = l < S0 && inside(x, y);
        long c = oL + g2(x, y);
#pragma unroll
        for (int g = 0; g < G; g++)
          pf_f[g][q] = (ok && g < ntr) ? fin0[(long)g * a.tstride + c] : TF(0);
      }
#pragma unroll
      for (int q = 0; q < C1; q++) {
        int l = tid + q * NT, x, y;
        rxy(R1, l, x, y);
        pf_ho[q] = (l < S1 && inside(x, y)) ? a.ho[oL + g2(x, y)] : 1.0;
      }
#pragma unroll
      for (int q = 0; q < CU1; q++) {
        int l = tid + q * NT, x, y;
        rxy(FU1, l, x, y);
        bool ok = l < SU1 && inside(x, y);
        pf_uu[q] = ok ? a.uu[oL + g2(x, y)] : 0.0;
        pf_hu[q] = ok ? a.hu[oL + g2(x, y)] : 1.0;
      }
#pragma unroll
      for (int q = 0; q < CV2; q++) {
        int l = tid + q * NT, x, y;
        rxy(FV2, l, x, y);
        bool ok = l < SV2 && inside(x, y);
        pf_vv[q] = ok ? a.vv[oL + g2(x, y)] : 0.0;
        pf_hv[q] = ok ? a.hv[oL + g2(x, y)] : 1.0;
      }
    }
    if (K >= 1 && K <= kmax) {
      const long oK = NXY * (long)K;
#pragma unroll
      for (int q = 0; q < C2; q++) {
        int l = tid + q * NT, x, y;
        rxy(R2, l, x, y);
        pf_ww[q] = (l < S2 && inside(x, y)) ? a.ww[oK + g2(x, y)] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < CU5; q++) {
        int l = tid + q * NT, x, y;
        rxy(FU5, l, x, y);
        bool ok = l < SU5 && inside(x, y);
        pf_uuK[q] = ok ? a.uu[oK + g2(x, y)] : 0.0;
        pf_huK[q] = ok ? a.hu[oK + g2(x, y)] : 1.0;
      }
#pragma unroll
      for (int q = 0; q < CV4; q++) {
        int l = tid + q * NT, x, y;
        rxy(FV4, l, x, y);
        bool ok = l < SV4 && inside(x, y);
        pf_vvK[q] = ok ? a.vv[oK + g2(x, y)] : 0.0;
        pf_hvK[q] = ok ? a.hv[oK + g2(x, y)] : 1.0;
      }
    }
  };
  auto stage = [&](int L, int K) {
    if (L >= 1 && L <= kmax) {
#pragma unroll
      for (int q = 0; q < C0; q++) {
        int l = tid + q * NT;
        if (l < S0)
#pragma unroll
          for (int g = 0; g < G; g++) F0[g][l] = pf_f[g][q];
      }
#pragma unroll
      for (int q = 0; q < C1; q++) {
        int l = tid + q * NT;
        if (l < S1) hoL[l] = pf_ho[q];
      }
#pragma unroll
      for (int q = 0; q < CU1; q++) {
        int l = tid + q * NT;
        if (l < SU1) { uuL[l] = pf_uu[q]; huL[l] = pf_hu[q]; }
      }
#pragma unroll
      for (int q = 0; q < CV2; q++) {
        int l = tid + q * NT;
        if (l < SV2) { vvL[l] = pf_vv[q]; hvL[l] = pf_hv[q]; }
      }
    }
    if (K >= 1 && K <= kmax) {
#pragma unroll
      for (int q = 0; q < C2; q++) {
        int l = tid + q * NT;
        if (l < S2) wwK[K & 1][l] = pf_ww[q];
      }
#pragma unroll
      for (int q = 0; q < CU5; q++) {
        int l = tid + q * NT;
        if (l < SU5) { uuK[l] = pf_uuK[q]; huK[l] = pf_huK[q]; }
      }
#pragma unroll
      for (int q = 0; q < CV4; q++) {
        int l = tid + q * NT;
        if (l < SV4) { vvK[l] = pf_vvK[q]; hvK[l] = pf_hvK[q]; }
      }
    }
  };

  prefetch(1, -1);
  // masks over R0
  for (int l = tid; l < S0; l += NT) {
    int x, y;
    rxy(R0, l, x, y);
    bool in = inside(x, y);
    long c = g2(x, y);
    m_uf[l] = in ? a.m_uflux[c] : 0;
    m_vf[l] = in ? a.m_vflux[c] : 0;
    m_uu[l] = in ? a.m_uupd[c] : 0;
    m_vu[l] = in ? a.m_vupd[c] : 0;
    m_az[l] = in ? (a.az[c] == 1) : 0;
  }
  // level 0: never advanced horizontally; seeds the rings of the w sub-step
  for (int l = tid; l < S2; l += NT) {
    int x, y;
    rxy(R2, l, x, y);
    bool in = inside(x, y);
    long c = g2(x, y);
    D2r[0][l] = in ? a.ho[c] : 1.0;
    wwK[0][l] = in ? a.ww[c] : 0.0;
    if constexpr (ADV) A2r[0][l] = 0.0;
    for (int g = 0; g < ntr; g++) F2r[0][g][l] = in ? fin0[(long)g * a.tstride + c] : TF(0);
  }
  // level 0 of the outputs
  for (int l = tid; l < ST; l += NT) {
    int x, y;
    rxy(RT, l, x, y);
    if (!inside(x, y)) continue;
    long c = g2(x, y);
    for (int g = 0; g < ntr; g++) fout0[(long)g * a.tstride + c] = fin0[(long)g * a.tstride + c];
    if (grp == 0 && a.hires) a.hires[c] = a.ho[c];
    if constexpr (ADV) {
      if (a.advres) a.advres[c] = 0.0;
    }
  }

  for (int m = 1; m <= kmax + 2; m++) {
    const int L = m, K = m - 2;
    const bool doL = L <= kmax, doK = K >= 1;
    const long oK = NXY * (long)K;
    const int sL = L % 4, sK = K % 4, sK1 = (K + 1) % 4;
    const int aL = L % 3, aK = (K + 3) % 3;

    __syncthreads();  // the previous level is done with the staging buffers
    stage(L, K);
    prefetch(L + 1, K + 1);
    __syncthreads();

    // ---- shared: hi through u/2 (level L) and w (level K)
    if (doL) {
      for (int l = tid; l < S1; l += NT) {
        int x, y;
        rxy(R1, l, x, y);
        double D0 = hoL[l], D1 = D0;
        if (inside(x, y) && x >= 1 + isub && x <= NX - 2 - isub && m_uu[R0.idx(x, y)])
          D1 = D0 - dth * (uuL[FU1.idx(x, y)] * dy - uuL[FU1.idx(x - 1, y)] * dy) * ard1;
        D1s[l] = D1;
      }
    }
    if (doK) {
      for (int l = tid; l < S2; l += NT) {
        int x, y;
        rxy(R2, l, x, y);
        double D2 = D2r[sK][l], D3 = D2;
        if (inside(x, y) && x <= NX - 2 && y <= NY - 2 && m_az[R0.idx(x, y)] &&
            K <= kmax - kshift)
          D3 = D2 - dt * (wwK[K & 1][l] - wwK[(K - 1) & 1][l]);
        D3s[l] = D3;
      }
    }
    __syncthreads();
    // ---- shared: hi through v/2 (level L) and the second v/2 (level K)
    if (doL) {
      for (int l = tid; l < S2; l += NT) {
        int x, y;
        rxy(R2, l, x, y);
        double D1 = D1s[R1.idx(x, y)], D2 = D1;
        if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)])
          D2 = D1 - dth * (vvL[FV2.idx(x, y)] * dx - vvL[FV2.idx(x, y - 1)] * dx) * ard1;
        D2r[sL][l] = D2;
      }
    }
    if (doK) {
      for (int l = tid; l < S4; l += NT) {
        int x, y;
        rxy(R4, l, x, y);
        double D3 = D3s[R2.idx(x, y)], D4 = D3;
        if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)])
          D4 = D3 - dth * (vvK[FV4.idx(x, y)] * dx - vvK[FV4.idx(x, y - 1)] * dx) * ard1;
        D4s[l] = D4;
      }
    }
    __syncthreads();

    for (int g = 0; g < ntr; g++) {
      TF* fout = fout0 + (long)g * a.tstride;
      const TF* F0g = F0[g];
      if (doL) {
        // u/2 fluxes on faces FU1 (adv_split_u.F90:219-265)
        for (int l = tid; l < SU1; l += NT) {
          int x, y;
          rxy(FU1, l, x, y);
          TF fl = TF(0);
          if (inside(x, y) && x >= isub && x <= NX - 2 - isub)
            fl = face_flux<TF, EXACT>(scheme, F0g, R0, x, y, 1, 0, m_uf, R0, uuL[l], huL[l],
                                      dth, dx, AH);
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < S1; l += NT) {
          int x, y;
          rxy(R1, l, x, y);
          TF f0 = F0g[R0.idx(x, y)], f1 = f0;
          double advn = 0.0;
          if (inside(x, y) && x >= 1 + isub && x <= NX - 2 - isub && m_uu[R0.idx(x, y)])
            f1 = cell_update<TF, EXACT>(hoL[l], D1s[l], f0, FL[FU1.idx(x, y)],
                                        FL[FU1.idx(x - 1, y)], 0.5, dy, ard1, dt, &advn);
          F1[l] = f1;
          if constexpr (ADV) A1s[l] = 0.0 + advn;
        }
        __syncthreads();
        // v/2 fluxes on faces FV2 (adv_split_v.F90:77-119)
        for (int l = tid; l < SV2; l += NT) {
          int x, y;
          rxy(FV2, l, x, y);
          TF fl = TF(0);
          if (inside(x, y) && y >= isub && y <= NY - 2 - isub)
            fl = face_flux<TF, EXACT>(scheme, F1, R1, x, y, 0, 1, m_vf, R0, vvL[l], hvL[l],
                                      dth, dy, AH);
          FL[l] = fl;
        }
        __syncthreads();
      }
      // v/2 update (level L) and w (level K) on R2; both touch column l only
      for (int l = tid; l < S2; l += NT) {
        int x, y;
        rxy(R2, l, x, y);
        if (doL) {
          int l1 = R1.idx(x, y);
          TF f1 = F1[l1], f2 = f1;
          double advn = 0.0;
          if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)])
            f2 = cell_update<TF, EXACT>(D1s[l1], D2r[sL][l], f1, FL[FV2.idx(x, y)],
                                        FL[FV2.idx(x, y - 1)], 0.5, dx, ard1, dt, &advn);
          F2r[sL][g][l] = f2;
          if constexpr (ADV) A2r[aL][l] = A1s[l1] + advn;
        }
        if (doK) {
          TF f2 = F2r[sK][g][l], f3 = f2;
          double advn = 0.0;
          if (inside(x, y) && x <= NX - 2 && y <= NY - 2 && m_az[R0.idx(x, y)]) {
            const double dti = dt;  // splitfac 1 for w
            TF wprev;
            if (K == 1) {
              wprev = kshift ? wflux_at<TF>(vscheme, 0, kshift, kmax, wwK[0][l], TF(0),
                                            F2r[0][g][l], F2r[1][g][l], F2r[2][g][l], D2r[0][l],
                                            D2r[1][l], dti)
                             : TF(0);
            } else {
              wprev = WF[g][l];
            }
            TF wk = TF(0);
            if (K <= kmax - 1) {
              TF fkm1 = F2r[(K + 3) % 4][g][l];
              TF fk1 = F2r[(K + 1) % 4][g][l];
              TF fk2 = (K + 2 <= kmax) ? F2r[(K + 2) % 4][g][l] : TF(0);
              wk = wflux_at<TF>(vscheme, K, kshift, kmax, wwK[K & 1][l], fkm1, f2, fk1, fk2,
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
      }
      if (doK) {
        __syncthreads();
        // second v/2 fluxes on FV4 from F3
        for (int l = tid; l < SV4; l += NT) {
          int x, y;
          rxy(FV4, l, x, y);
          TF fl = TF(0);
          if (inside(x, y) && y >= isub && y <= NY - 2 - isub)
            fl = face_flux<TF, EXACT>(scheme, F3, R2, x, y, 0, 1, m_vf, R0, vvK[l], hvK[l],
                                      dth, dy, AH);
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < S4; l += NT) {
          int x, y;
          rxy(R4, l, x, y);
          int l2 = R2.idx(x, y);
          TF f3 = F3[l2], f4 = f3;
          double advn = 0.0;
          if (inside(x, y) && y >= 1 + isub && y <= NY - 2 - isub && m_vu[R0.idx(x, y)])
            f4 = cell_update<TF, EXACT>(D3s[l2], D4s[l], f3, FL[FV4.idx(x, y)],
                                        FL[FV4.idx(x, y - 1)], 0.5, dx, ard1, dt, &advn);
          F4[l] = f4;
          if constexpr (ADV) A4s[l] = A3s[l2] + advn;
        }
        __syncthreads();
        // second u/2 fluxes on FU5 from F4
        for (int l = tid; l < SU5; l += NT) {
          int x, y;
          rxy(FU5, l, x, y);
          TF fl = TF(0);
          if (inside(x, y) && x >= isub && x <= NX - 2 - isub)
            fl = face_flux<TF, EXACT>(scheme, F4, R4, x, y, 1, 0, m_uf, R0, uuK[l], huK[l],
                                      dth, dx, AH);
          FL[l] = fl;
        }
        __syncthreads();
        for (int l = tid; l < ST; l += NT) {
          int x, y;
          rxy(RT, l, x, y);
          if (!inside(x, y)) continue;
          int l4 = R4.idx(x, y);
          TF f4 = F4[l4], f5 = f4;
          double D4 = D4s[l4], D5 = D4;
          double advn = 0.0;
          long c = oK + g2(x, y);
          if (x >= 1 + isub && x <= NX - 2 - isub && m_uu[R0.idx(x, y)]) {
            D5 = D4 - dth * (uuK[FU5.idx(x, y)] * dy - uuK[FU5.idx(x - 1, y)] * dy) * ard1;
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
      __syncthreads();  // FL and F1/F3/F4 are reused by the next tracer
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
      return adv ? launch_one<double, 1, true, true, 64, 4, 512>(st, a)
                 : launch_one<double, 1, true, false, 64, 4, 512>(st, a);
    case 2:
      return adv ? launch_one<double, 1, true, true, 16, 8, 256>(st, a)
                 : launch_one<double, 1, true, false, 16, 8, 256>(st, a);
    default:
      return adv ? launch_one<double, 1, true, true, 32, 8, 512>(st, a)
                 : launch_one<double, 1, true, false, 32, 8, 512>(st, a);
  }
}

hipError_t launch_fused_f32(hipStream_t st, const FusedArgs<float>& a, int tile, int group) {
#define GT_F32(G_, TX_, TY_) launch_one<float, G_, false, false, TX_, TY_, (TX_ * TY_ >= 256 ? 512 : 256)>(st, a)
  switch (group) {
    case 1:
      return tile == 1 ? GT_F32(1, 64, 4) : tile == 2 ? GT_F32(1, 16, 8) : GT_F32(1, 32, 8);
    case 2:
      return tile == 1 ? GT_F32(2, 64, 4) : tile == 2 ? GT_F32(2, 16, 8) : GT_F32(2, 32, 8);
    case 6:  // 64x4 with six tracers does not fit the 160 KB of LDS
      return tile == 1 ? GT_F32(3, 64, 4) : tile == 2 ? GT_F32(6, 16, 8) : GT_F32(6, 32, 8);
    default:
      return tile == 1 ? GT_F32(3, 64, 4) : tile == 2 ? GT_F32(3, 16, 8) : GT_F32(3, 32, 8);
  }
#undef GT_F32
}

}  // namespace gt
