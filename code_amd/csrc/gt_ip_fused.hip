// gt_ip_fused.hip -- the equation of state (linear, UNESCO 1980 or Jackett
// et al.; src/3d/eqstate.F90:94-261) fused into the internal pressure
// gradients of Blumberg-Mellor (src/3d/ip_blumberg_mellor.F90) and
// Shchepetkin-McWilliams (src/3d/ip_shchepetkin_mcwilliams.F90), either or
// both, in two launches:
//   k_ip_columns  one thread per column: the level-0 zeros, the interface
//                 heights of Blumberg-Mellor and the layer centres of
//                 Shchepetkin-McWilliams, both from one read of hn;
//   k_ip_rows     one wavefront per row of a 60-column strip, marching down
//                 the levels: the buoyancy formed from T and S where it is
//                 read (never stored), Blumberg-Mellor's two downward
//                 recurrences, the pressure integral P of
//                 Shchepetkin-McWilliams and the Jacobians of both faces.
// Every expression is the per-method kernels' (gt_kernels.hip,
// gt_eqstate.hip), in their order, so the gradients are bit for bit those of
// gt_eos_linear or gt_do_eqstate followed by gt_ip_blumberg_mellor /
// gt_ip_shchepetkin_mcwilliams.
#include <hip/hip_runtime.h>

#include "gt_eos.h"
#include "gt_ip.h"
#include "gt_kernels.h"

namespace gt {

template <bool BM, bool SM>
__global__ __launch_bounds__(256) void k_ip_columns(Dims d, const int32_t* __restrict__ az,
                                                    const double* __restrict__ H,
                                                    const double* __restrict__ hn,
                                                    double* __restrict__ zzb,
                                                    double* __restrict__ zzs,
                                                    double* __restrict__ ixb,
                                                    double* __restrict__ iyb,
                                                    double* __restrict__ ixs,
                                                    double* __restrict__ iys) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  const int jj = blockIdx.y * blockDim.y + threadIdx.y;
  if (ii >= d.NX || jj >= d.NY) return;
  const long NXY = d.NXY;
  const int kmax = d.kmax;
  const long c0 = ii + (long)d.NX * jj;
  if constexpr (BM) {
    zzb[c0] = 0.0;
    ixb[c0] = 0.0;
    iyb[c0] = 0.0;
  }
  if constexpr (SM) {
    zzs[c0] = 0.0;
    ixs[c0] = 0.0;
    iys[c0] = 0.0;
  }
  if (az[c0] < 1) return;
  // Blumberg-Mellor's interfaces over [imin..imax+1]x[jmin..jmax+1]
  // (ip_blumberg_mellor.F90:83-92), Shchepetkin-McWilliams' centres over
  // the whole halo range (ip_shchepetkin_mcwilliams.F90:90-99)
  const bool bm = BM && ii >= 2 && ii <= d.NX - 2 && jj >= 2 && jj <= d.NY - 2;
  if (!bm && !SM) return;
  const double h1 = hn[c0 + NXY];
  double zb = -H[c0] + h1;
  double zs = -H[c0] + 0.5 * h1;
  if (bm) zzb[c0 + NXY] = zb;
  if constexpr (SM) zzs[c0 + NXY] = zs;
  double hprev = h1;
  constexpr int kB = 8;  // levels whose heights are loaded together
  for (int kb = 2; kb <= kmax; kb += kB) {
    double q[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) q[u] = hn[c0 + NXY * min(kb + u, kmax)];
#pragma unroll
    for (int u = 0; u < kB; u++) {
      const bool on = kb + u <= kmax;
      const double hk = q[u];
      const double zbn = zb + hk;
      const double zsn = zs + 0.5 * (hprev + hk);
      zb = on ? zbn : zb;
      zs = on ? zsn : zs;
      hprev = on ? hk : hprev;
      if (on) {
        if (bm) zzb[c0 + NXY * (kb + u)] = zb;
        if constexpr (SM) zzs[c0 + NXY * (kb + u)] = zs;
      }
    }
  }
}

// the downward recurrence of Shchepetkin-McWilliams' pressure integral P of
// one column (ip_shchepetkin_mcwilliams.F90:108-146): the state at level k
// holds b1 = b(k+1), b0 = b(k), z1 = z(k+1), z0 = z(k), the harmonic means of
// the differences above, and p = P(k+1) until step() forms P(k)
struct PCol {
  double p, hR1, hZ1, rk, zk, b1, b0, z1, z0;
  // P(kmax) from the top two levels
  __device__ __forceinline__ void top(double bt, double bb, double zt, double zb, double ht) {
    b1 = bt; b0 = bb; z1 = zt; z0 = zb;
    const double cff1 = 1.0 / (z1 - z0);
    const double cff2 = 0.5 * (b1 - b0) * 0.5 * ht * cff1;
    p = (b1 + cff2) * 0.5 * ht;
    rk = b1 - b0;
    zk = z1 - z0;
    hR1 = sm_hmR(rk, rk);
    hZ1 = sm_hmZ(zk, zk);
  }
  // P(k) for k < kmax from b(k-1), z(k-1) (unused at k = 1, where the
  // differences below are those above: dR(0) = dR(1)), then the window moves
  // down one level
  __device__ __forceinline__ void step(int k, double bm, double zm) {
    double rkm, zkm;
    if (k >= 2) {
      rkm = b0 - bm;
      zkm = z0 - zm;
    } else {
      rkm = b1 - b0;
      zkm = z1 - z0;
      bm = 0.0;
      zm = 0.0;
    }
    const double hR = sm_hmR(rk, rkm), hZ = sm_hmZ(zk, zkm);
    p = p + 0.5 * ((b1 + b0) * (z1 - z0) -
                   SM_ONEFIFTH * ((hR1 - hR) * (z1 - z0 - SM_C12 * (hZ1 + hZ)) -
                                  (hZ1 - hZ) * (b1 - b0 - SM_C12 * (hR1 + hR))));
    hR1 = hR;
    hZ1 = hZ;
    rk = rkm;
    zk = zkm;
    b1 = b0;
    b0 = bm;
    z1 = z0;
    z0 = zm;
  }
};

// the face term of Shchepetkin-McWilliams (ip_shchepetkin_mcwilliams.F90:
// 188-209 for x, 262-284 for y): b, z at the two cells of the face, the
// harmonic means either side, P either side, the surface elevation either
// side and b(kmax) either side
__device__ __forceinline__ double sm_face(double b0, double b1, double z0, double z1,
                                          double dR0, double dR1, double dZ0, double dZ1,
                                          double p0, double p1, double h, double dm1, double s0,
                                          double s1, double k0, double k1) {
  double AJ = p1 - p0;
  const double FC = 0.5 * ((b1 + b0) * (z1 - z0) -
                           SM_ONEFIFTH * ((dR1 - dR0) * (z1 - z0 - SM_C12 * (dZ1 + dZ0)) -
                                          (dZ1 - dZ0) * (b1 - b0 - SM_C12 * (dR1 + dR0))));
  AJ = AJ + FC;
  return AJ * dm1 * h - h * dm1 * (s1 - s0) * 0.5 * (k1 + k0);
}

// levels by which k_ip_rows loads ahead of its arithmetic.  The kernel must
// stay at 4 wavefronts per SIMD (<= 128 VGPRs), so that all workgroups of the
// headline grid are resident at once: the linear EOS fits one level of
// lookahead (128 VGPRs with hun and hvn loaded where a level starts and the
// neighbours' face constants read from LDS; 175 against 181 us for both
// methods at 400x400x40), the full equations of state do not (132 VGPRs, 3
// wavefronts per SIMD, two rounds of workgroups) and load each level where
// it starts, the other wavefronts hiding the latency
template <int EOS>
constexpr int kIpAhead = EOS == 1 ? 1 : 0;

// the operands of one level k of the march that come from memory, loaded
// kIpAhead<EOS> levels ahead of their use (raw T and S, the buoyancy formed where
// it is used, so with a lookahead the loads stay in flight across the
// arithmetic of the levels before).  Row y is the wavefront's own row; the
// rows around it come from the other wavefronts of the workgroup through LDS.
struct IpLevel {
  double t0, s0, z0;  // T S zz(centres) at (y, k-1) (SM: the P step); T S at (y, k) (BM only)
  double te, se, ze;  // T S zz(centres) at (y-1, k) (first wavefront) or (y+1, k) (carrier);
                      // BM only: T S zz(interfaces) at (y+1, k) (last wavefront)
  double zb0;         // zz(interfaces) at (y, k) (BM)
};
// hun and hvn at (y, k): loaded where level k starts, before the next
// level's operands, so they hold no registers across a level
struct IpH {
  double hu, hv;
};

// One workgroup of W+1 wavefronts (W without Shchepetkin-McWilliams) per
// strip of 60 interior columns x W rows, marching down the levels;
// wavefront w takes row y = j0 + w and lane l column x = 60*sx + 1 + l
// (lanes 1..60 of wavefronts 0..W-1 write; the last wavefront only carries
// row j0+W for the one below it).  The i-neighbours x-1, x+1, x+2 are
// adjacent lanes (DPP); the j-neighbours are the other wavefronts' rows,
// exchanged through LDS at each level, so each wavefront loads its own row
// only (the first also row j0-1, the last row j0+W+1, for the harmonic means
// across y at the workgroup's edges).  Per level: each wavefront takes the P
// step of its row and publishes b, z (and Blumberg-Mellor's interface
// height); after a barrier it forms the two harmonic means across y of its
// row and publishes them with P; after a second barrier it forms its x face
// (means shared with the next lane), its y face (means of row y+1 from the
// wavefront above) and Blumberg-Mellor's recurrences.  Every mean is formed
// once and rounds as the per-face formula.  Rows past NY-1 and columns past
// NX-1 are clamped (read, never written).
// EOS 1: the linear equation of state at every cell; 2 (UNESCO 1980) and 3
// (Jackett et al. at p = 0): the density of water columns from T and S
// (S < 0 taken as 0 when nonneg), and on land columns (az = 0) the rho the
// caller holds, which do_eqstate leaves there and turns into buoyancy
// (eqstate.F90:154-158,198-222,252-261): Shchepetkin-McWilliams' harmonic
// means reach one column past a water face.
template <int EOS, bool BM, bool SM, int W>
__global__ __launch_bounds__(64 * (W + 1)) void k_ip_rows(
    Dims d, int nsx, int nby, const int32_t* __restrict__ au, const int32_t* __restrict__ av,
    const int32_t* __restrict__ az, const double* __restrict__ rho, int nonneg,
    const double* __restrict__ T, const double* __restrict__ S, double T0, double S0,
    double dtr0, double dsr0, const double* __restrict__ hn, const double* __restrict__ hun,
    const double* __restrict__ hvn, const double* __restrict__ zzb,
    const double* __restrict__ zzs, const double* __restrict__ sseo, double* __restrict__ ixb,
    double* __restrict__ iyb, double* __restrict__ ixs, double* __restrict__ iys, double dxm1,
    double dym1) {
  // workgroups are dealt round-robin to the 8 XCDs: workgroup b takes the
  // work of (b%8)*per + b/8, so each XCD marches a contiguous run of strips
  // (rows fastest) whose shared rows meet in its L2; the test is uniform
  // over the workgroup, so every wavefront that starts reaches each barrier
  constexpr int NW = W + (SM ? 1 : 0);
  const long nblk = (long)nsx * nby;
  const long per = (nblk + 7) / 8;
  const long blk = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (blk >= nblk) return;
  const int by = (int)(blk % nby), sx = (int)(blk / nby);
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  const bool carrier = SM && wave == W;  // row j0+W: P and means only
  const bool first = wave == 0, last = wave == NW - 1;
  const int NX = d.NX, NY = d.NY, kmax = d.kmax;
  const long NXY = d.NXY;
  const int x = 60 * sx + 1 + lane;  // >= 1
  const int xc = x < NX ? x : NX - 1;
  const int y = 2 + W * by + wave;
  // 32-bit cell offsets within a level (a level holds < 2^31 cells)
  auto cell = [&](int yy) -> unsigned { return (unsigned)(xc + NX * (yy < NY ? yy : NY - 1)); };
  const unsigned c0 = cell(y), cp = cell(y + 1);
  // the extra row of the edge wavefronts (SM): y-1 for the first, y+1 for
  // the carrier (a workgroup of one row has both: the first is not the carrier)
  const unsigned ce = carrier ? cp : cell(y - 1);
  const bool out = !carrier && lane >= 1 && lane <= 60 && x <= NX - 3 && y <= NY - 3;
  const bool ux = out && au[c0] >= 1, vy = out && av[c0] >= 1;
  // water columns of c0, cp and the extra row's cell (te, se below)
  const bool w0 = EOS == 1 || az[c0] > 0, wp = EOS == 1 || az[cp] > 0;
  const bool wte = SM ? EOS == 1 || az[ce] > 0 : wp;
  const bool nn = nonneg != 0;
  // the buoyancy of a cell from its T and S, or (land, full equations of
  // state) from its rho, which the loads put in place of T
  auto buoy = [&](double t, double s, bool wet) -> double {
    if constexpr (EOS == 1) {
      return eos_linear_buoy(t, s, T0, S0, dtr0, dsr0);
    } else {
      return buoy_of_rho(wet ? eqstate_rho<EOS>(t, s, nn) : t);
    }
  };
  auto ldt = [&](long o, unsigned c, bool wet) { return wet ? (T + o)[c] : (rho + o)[c]; };
  auto lds = [&](long o, unsigned c, bool wet) { return wet ? (S + o)[c] : 0.0; };
  auto load = [&](int k, IpLevel& L) {
    const long ok = NXY * (long)k, okm = NXY * (long)(k >= 2 ? k - 1 : 1);
    const long o0 = SM ? okm : ok;
    L.t0 = ldt(o0, c0, w0);
    L.s0 = lds(o0, c0, w0);
    if constexpr (SM) {
      L.z0 = (zzs + o0)[c0];
      if (first || carrier) {
        L.te = ldt(ok, ce, wte);
        L.se = lds(ok, ce, wte);
        L.ze = (zzs + ok)[ce];
      }
    } else if (last) {
      // Blumberg-Mellor alone: the last wavefront reads row y+1 itself
      L.te = ldt(ok, cp, wte);
      L.se = lds(ok, cp, wte);
      L.ze = (zzb + ok)[cp];
    }
    if constexpr (BM) L.zb0 = (zzb + ok)[c0];
  };
  auto load_h = [&](int k, IpH& H) {
    const long ok = NXY * (long)k;
    if (!carrier) {
      H.hu = (hun + ok)[c0];
      H.hv = (hvn + ok)[c0];
    }
  };
  // LDS: per wavefront and lane, b, z and zz(interfaces) of its row (slot A),
  // P and the two means across y (slot B); two of each by level parity
  __shared__ double shA[2][3][NW][64];
  __shared__ double shB[SM ? 2 : 1][3][NW][64];
  // Shchepetkin-McWilliams' face constants of each row: buoyancy at kmax and
  // sseo, read by the lane below (x face) and the wavefront below (y face)
  __shared__ double shK[2][NW][64];
  const int lu = lane < 63 ? lane + 1 : 63;

  // ---- level kmax
  const long okx = NXY * (long)kmax;
  const double kb0 = buoy(ldt(okx, c0, w0), lds(okx, c0, w0), w0);
  // Shchepetkin-McWilliams: P of the row, the face constants
  PCol ps{};
  double ss0 = 0;
  // Blumberg-Mellor (ip_blumberg_mellor.F90:99-149), as k_bm_idpd
  double grdlx = 0, buoylx = 0, hkx = 0, prgrx = 0, grdly = 0, buoyly = 0, hky = 0, prgry = 0;

  // the exchange and the faces of level k, from the row's b(k), z(k) (SM:
  // the P state, P(k) in ps.p), zz(interfaces)(k) and hun, hvn (L)
  auto level_faces = [&](int k, int slot, double bk, double zk, const IpLevel& L, const IpH& H,
                         bool top) {
    const long ok = NXY * (long)k;
    shA[slot][0][wave][lane] = bk;
    if constexpr (SM) shA[slot][1][wave][lane] = zk;
    if constexpr (BM) shA[slot][2][wave][lane] = L.zb0;
    __syncthreads();
    double bn = 0, zn = 0, zbn = 0;  // row y+1 at k
    if (!SM && last) {
      bn = buoy(L.te, L.se, wte);
      zbn = L.ze;
    } else if (!carrier) {
      bn = shA[slot][0][wave + 1][lane];
      if constexpr (SM) zn = shA[slot][1][wave + 1][lane];
      if constexpr (BM) zbn = shA[slot][2][wave + 1][lane];
    }
    if constexpr (SM) {
      // the means of rows y-1|y|y+1, shared with the wavefront below
      double bm, zm, bp, zp;
      if (first) {
        bm = buoy(L.te, L.se, wte);
        zm = L.ze;
      } else {
        bm = shA[slot][0][wave - 1][lane];
        zm = shA[slot][1][wave - 1][lane];
      }
      if (carrier) {
        bp = buoy(L.te, L.se, wte);
        zp = L.ze;
      } else {
        bp = bn;
        zp = zn;
      }
      const double hZ = sm_hmx(zk - zm, zp - zk), hR = sm_hmx(bk - bm, bp - bk);
      shB[slot][0][wave][lane] = ps.p;
      shB[slot][1][wave][lane] = hZ;
      shB[slot][2][wave][lane] = hR;
      if (!carrier) {
        // x: the means of columns x-1|x|x+1 on this lane, x|x+1|x+2 on the next
        const double p0 = ps.p;
        const double bL = from_lower_lane(bk), zL = from_lower_lane(zk);
        const double bU = from_upper_lane(bk), zU = from_upper_lane(zk);
        const double pU = from_upper_lane(p0);
        const double dZ = sm_hmx(zk - zL, zU - zk), dR = sm_hmx(bk - bL, bU - bk);
        const double dZ1 = from_upper_lane(dZ), dR1 = from_upper_lane(dR);
        if (ux)
          (ixs + ok)[c0] = sm_face(bk, bU, zk, zU, dR, dR1, dZ, dZ1, p0, pU, H.hu, dxm1, ss0,
                                   shK[1][wave][lu], kb0, shK[0][wave][lu]);
      }
      __syncthreads();
      if (!carrier) {
        const double pP = shB[slot][0][wave + 1][lane];
        const double hZ1 = shB[slot][1][wave + 1][lane];
        const double hR1 = shB[slot][2][wave + 1][lane];
        if (vy)
          (iys + ok)[c0] = sm_face(bk, bn, zk, zn, hR, hR1, hZ, hZ1, ps.p, pP, H.hv, dym1, ss0,
                                   shK[1][wave + 1][lane], kb0, shK[0][wave + 1][lane]);
      }
    }
    if constexpr (BM) {
      if (!carrier) {
        const double b0 = bk, be = from_upper_lane(bk);
        const double z0 = L.zb0, ze = from_upper_lane(L.zb0);
        if (top) {
          // level kmax starts the recurrences
          grdlx = (be - b0) * dxm1;
          buoylx = 0.5 * (be + b0);
          hkx = H.hu;
          prgrx = grdlx * 0.5 * hkx;
          grdly = (bn - b0) * dym1;
          buoyly = 0.5 * (bn + b0);
          hky = H.hv;
          prgry = grdly * 0.5 * hky;
          if (ux) (ixb + ok)[c0] = hkx * prgrx;
          if (vy) (iyb + ok)[c0] = hky * prgry;
        } else {
          {
            const double grdu = grdlx;
            grdlx = (be - b0) * dxm1;
            const double buoyu = buoylx;
            buoylx = 0.5 * (be + b0);
            const double dxz = (ze - z0) * dxm1;
            const double hkk = H.hu;
            prgrx = prgrx + 0.5 * (grdu + grdlx) * 0.5 * (hkk + hkx) - dxz * (buoyu - buoylx);
            if (ux) (ixb + ok)[c0] = hkk * prgrx;
            hkx = hkk;
          }
          {
            const double grdu = grdly;
            grdly = (bn - b0) * dym1;
            const double buoyu = buoyly;
            buoyly = 0.5 * (bn + b0);
            const double dyz = (zbn - z0) * dym1;
            const double hkk = H.hv;
            prgry = prgry + 0.5 * (grdu + grdly) * 0.5 * (hkk + hky) - dyz * (buoyu - buoyly);
            if (vy) (iyb + ok)[c0] = hkk * prgry;
            hky = hkk;
          }
        }
      }
    }
  };

  {
    IpLevel top;
    IpH hh;
    load(kmax, top);
    load_h(kmax, hh);
    if constexpr (SM) {
      ps.top(kb0, buoy(top.t0, top.s0, w0), (zzs + okx)[c0], top.z0, (hn + okx)[c0]);
      ss0 = sseo[c0];
      shK[0][wave][lane] = kb0;
      shK[1][wave][lane] = ss0;
      level_faces(kmax, kmax & 1, ps.b1, ps.z1, top, hh, true);
    } else {
      level_faces(kmax, kmax & 1, kb0, 0.0, top, hh, true);
    }
  }

  // ---- levels kmax-1 .. 1
  auto level = [&](int k, const IpLevel& L, const IpH& H) {
    if constexpr (SM) {
      ps.step(k, buoy(L.t0, L.s0, w0), L.z0);
      level_faces(k, k & 1, ps.b1, ps.z1, L, H, false);
    } else {
      level_faces(k, k & 1, buoy(L.t0, L.s0, w0), 0.0, L, H, false);
    }
  };
  // a ring of NB operand buffers: level k's operands are loaded NB-1 levels
  // ahead (levels below 1 re-read level 1, harmlessly)
  constexpr int NB = kIpAhead<EOS> + 1;
  auto lv = [](int k) { return k >= 1 ? k : 1; };
  IpLevel ring[NB];
#pragma unroll
  for (int q = 0; q < NB - 1; q++) load(lv(kmax - 1 - q), ring[q]);
  for (int k0 = kmax - 1; k0 >= 1; k0 -= NB) {
#pragma unroll
    for (int q = 0; q < NB; q++) {
      const int k = k0 - q;
      if (k < 1) break;
      IpH hh;
      load_h(k, hh);
      load(lv(k - (NB - 1)), ring[(q + NB - 1) % NB]);
      level(k, ring[q], hh);
    }
  }
}

constexpr int kIpRowsW = 7;  // output rows per workgroup of k_ip_rows (W+1 wavefronts)

hipError_t launch_eos_ip_fused(hipStream_t st, Dims d, int eos, bool nonneg, bool bm, bool sm,
                               const int32_t* az, const int32_t* au, const int32_t* av,
                               const double* T, const double* S, const double* rho, double T0,
                               double S0, double dtr0, double dsr0,
                               const double* H, const double* hn, const double* hun,
                               const double* hvn, const double* sseo, double* zzb, double* ixb,
                               double* iyb, double* zzs, double* ixs, double* iys, double dx,
                               double dy) {
  const dim3 blk(64, 4, 1);
  const dim3 g1((d.NX + 63) / 64, (d.NY + 3) / 4, 1);
  constexpr int W = kIpRowsW;
  const int nsx = (d.NX - 4 + 59) / 60, nby = (d.NY - 4 + W - 1) / W;
  const long nblk = (long)nsx * nby;
  const dim3 g2((unsigned)(((nblk + 7) / 8) * 8));
  const double dxm1 = 1.0 / dx, dym1 = 1.0 / dy;
  const int nn = nonneg ? 1 : 0;
#define GT_IPF(E, B, M)                                                                       \
  do {                                                                                        \
    hipLaunchKernelGGL((k_ip_columns<B, M>), g1, blk, 0, st, d, az, H, hn, zzb, zzs, ixb, iyb, \
                       ixs, iys);                                                             \
    hipError_t e_ = hipGetLastError();                                                        \
    if (e_ != hipSuccess) return e_;                                                          \
    if (d.NX > 4 && d.NY > 4)                                                                 \
      hipLaunchKernelGGL((k_ip_rows<E, B, M, W>), g2, dim3(64 * (W + (M ? 1 : 0))), 0, st, d, \
                         nsx, nby, au, av, az, rho, nn, T, S, T0, S0, dtr0, dsr0, hn, hun,    \
                         hvn, zzb, zzs, sseo, ixb, iyb, ixs, iys, dxm1, dym1);                \
    return hipGetLastError();                                                                 \
  } while (0)
#define GT_IPF_EOS(E)                  \
  do {                                 \
    if (bm && sm) GT_IPF(E, true, true); \
    if (bm) GT_IPF(E, true, false);    \
    if (sm) GT_IPF(E, false, true);    \
    return hipSuccess;                 \
  } while (0)
  if (eos == 2) GT_IPF_EOS(2);
  if (eos == 3) GT_IPF_EOS(3);
  GT_IPF_EOS(1);
#undef GT_IPF_EOS
#undef GT_IPF
}

}  // namespace gt
