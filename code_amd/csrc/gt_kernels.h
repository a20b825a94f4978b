// gt_kernels.h -- host launchers for the transport kernels.
#pragma once

#include "gt_device.h"

namespace gt {

struct TileDesc {
  double* f;
  int ni, nj;
};

// the tiles of one exchange, passed by value as a kernel argument (no
// device-side table, so the exchange allocates nothing)
constexpr int kMaxTiles = 64;
struct TileSet {
  TileDesc t[kMaxTiles];
  long ring_off[kMaxTiles + 1];  // start of each tile's ghost ring in the flat index
};

// every launcher enqueues on `st` and returns the launch status
hipError_t launch_split_u(hipStream_t st, Dims d, int nslice, const double* f,
                          const double* fi_src, double* fi_dst, double* Di, double* adv,
                          const double* U, const double* DU, const uint8_t* mflux,
                          const uint8_t* mupd, double dt, double splitfac, int scheme,
                          double AH, double dx, double dy, double ard1);
hipError_t launch_split_v(hipStream_t st, Dims d, int nslice, const double* f,
                          const double* fi_src, double* fi_dst, double* Di, double* adv,
                          const double* V, const double* DV, const uint8_t* mflux,
                          const uint8_t* mupd, double dt, double splitfac, int scheme,
                          double AH, double dx, double dy, double ard1);
hipError_t launch_split_w(hipStream_t st, Dims d, const double* f, const double* fi_src,
                          double* fi_dst, double* hi, double* adv, const double* ww,
                          const int32_t* az, double dt, double splitfac, int scheme,
                          int kshift);
hipError_t launch_split_w_iter(hipStream_t st, Dims d, const double* f, const double* fi_src,
                               double* fi_dst, double* hi, double* adv, const double* ww,
                               const int32_t* az, double* scratch7, unsigned long long* nwarn,
                               double dt, double splitfac, int scheme, int kshift,
                               int itersmax);
hipError_t launch_finalise(hipStream_t st, Dims d, int k0, int nk, double* f, const double* ho,
                           const double* adv, const double* hi, const uint8_t* mfin,
                           double dt);
hipError_t launch_upstream_2dh(hipStream_t st, Dims d, int nslice, const double* f,
                               const double* fi_src, double* fi_dst, double* Di, double* adv,
                               const double* U, const double* V, const double* Dn,
                               const int32_t* az, double dt, double AH, double dx, double dy,
                               double ard1);
hipError_t launch_eos_linear(hipStream_t st, Dims d, const double* T, const double* S,
                             double* rho, double* buoy, double T0, double S0, double dtr0,
                             double dsr0);
// do_eqstate (eqstate.F90:94-271): method 1 linear (+ alpha/beta fill), 2
// UNESCO 1980, 3 Jackett et al. (2006); alpha/beta may be NULL
hipError_t launch_eqstate(hipStream_t st, Dims d, int method, bool nonnegsalt,
                          const int32_t* az, const double* T, const double* S,
                          const double* hn, double* rho, double* buoy, double* alpha,
                          double* beta, double T0, double S0, double dtr0, double dsr0);
// fn 0: rho_from_theta_unesco80(th=T, s=S) -> o1; 1: rho_from_theta -> o1 =
// dens0, o2 = densp; 2: eosall_from_theta -> o1 = rho_s, o2 = rho_th
hipError_t launch_eos_points(hipStream_t st, int fn, long n, const double* s, const double* th,
                             const double* p, double* o1, double* o2, bool nonnegsalt);
// input producers (gt_inputs.hip)
hipError_t launch_start_macro(hipStream_t st, Dims d, double split, double min_depth,
                              const double* H, const double* HU, const double* HV,
                              const double* z, double* sseo, double* ssen, double* ssuo,
                              double* ssun, double* ssvo, double* ssvn, double* Dn, double* Dun,
                              double* Dvn, double* Uint, double* Vint);
hipError_t launch_sigma(hipStream_t st, Dims d, bool equiv, double kmaxm1, const double* dga,
                        const double* H, const double* HU, const double* HV, const double* sseo,
                        const double* ssen, const double* ssuo, const double* ssun,
                        const double* ssvo, const double* ssvn, double* ho, double* hn,
                        double* huo, double* hun, double* hvo, double* hvn, double* zc);
hipError_t launch_gvc_init(hipStream_t st, Dims d, int kk, double d_gamma, double min_depth,
                           const double* lev, const double* H, const double* HU,
                           const double* HV, const double* sseo, const double* ssuo,
                           const double* ssvo, double* gga, double* hn, double* huo,
                           double* hun, double* hvo, double* hvn, int* bad);
hipError_t launch_gvc_step(hipStream_t st, Dims d, double relax_dt, double maxdepth,
                           double d_gamma, const double* gga, const double* H, const double* HU,
                           const double* HV, const double* ssen, const double* ssun,
                           const double* ssvn, double* ho, double* hn, double* huo, double* hun,
                           double* hvo, double* hvn, double* zc);
hipError_t launch_zc(hipStream_t st, Dims d, const double* H, const double* hn, double* zc);
hipError_t launch_ww_momentum(hipStream_t st, Dims d, double dt, double dx, double dy,
                              double ard1, const int32_t* az, const double* uu, const double* vv,
                              const double* ho, const double* hn, double* ww);
// vertical diffusion (gt_diffusion.hip); work holds 3 fields of N3
hipError_t launch_vertical_diffusion(hipStream_t st, Dims d, bool temperature, double dt,
                                     double cnpar, const int32_t* az, const double* hn,
                                     const double* nuh, double* f, double* work,
                                     const double* swr, const double* shf, const double* A,
                                     const double* g1, const double* g2, const double* dry_z,
                                     const double* S, double refl_frac, double min_bot_frac,
                                     double* rad);
// n systems, level-major; work holds 2*n*(lt+1)
hipError_t launch_tridiagonal(hipStream_t st, long n, int fi, int lt, const double* au,
                              const double* bu, const double* cu, const double* du,
                              double* value, double* work);
// uv_advect_3d averaging (gt_momentum.hip); work holds 6 fields of N3
hipError_t launch_uv_prep(hipStream_t st, Dims d, bool upart, const double* uu, const double* vv,
                          const double* ww, const double* ho, const double* hn,
                          const double* hun, const double* hvn, double* work);
hipError_t launch_ip_bm(hipStream_t st, Dims d, const int32_t* az, const int32_t* au,
                        const int32_t* av, const double* H, const double* hn,
                        const double* hun, const double* hvn, const double* buoy, double* zz,
                        double* idpdx, double* idpdy, double dx, double dy);
// the linear EOS fused into Blumberg-Mellor (bm) and/or Shchepetkin-McWilliams
// (sm): two launches, gradients bit for bit those of launch_eos_linear +
// launch_ip_bm / launch_ip_smcw (gt_ip_fused.hip)
// eos: do_eqstate's method (1 linear, 2 UNESCO 1980, 3 Jackett et al.); rho
// is read on land columns under methods 2 and 3 only
hipError_t launch_eos_ip_fused(hipStream_t st, Dims d, int eos, bool nonneg, bool bm, bool sm,
                               const int32_t* az, const int32_t* au, const int32_t* av,
                               const double* T, const double* S, const double* rho, double T0,
                               double S0, double dtr0, double dsr0,
                               const double* H, const double* hn, const double* hun,
                               const double* hvn, const double* sseo, double* zzb, double* ixb,
                               double* iyb, double* zzs, double* ixs, double* iys, double dx,
                               double dy);
hipError_t launch_ip_smcw(hipStream_t st, Dims d, const int32_t* az, const int32_t* au,
                          const int32_t* av, const double* H, const double* hn,
                          const double* hun, const double* hvn, const double* buoy,
                          const double* sseo, double* zz, double* P, double* idpdx,
                          double* idpdy, double dx, double dy);
hipError_t launch_mirror_3d(hipStream_t st, Dims d, double* f1, const double* f2);
hipError_t launch_mirror_2d(hipStream_t st, Dims d, double* f1, const double* f2);
hipError_t launch_tile_exchange(hipStream_t st, int ntiles, int px, const TileSet& tiles, int nk);
// SUBSTR_INI_PRESS (internal_pressure.F90:235-258) over the interior, all levels:
// first call keeps the gradients in x0/y0 and zeroes them, later calls subtract
hipError_t launch_substr_ini_press(hipStream_t st, Dims d, int first, double* idpdx,
                                   double* idpdy, double* x0, double* y0);
// partial must hold 4*nblocks doubles; out two doubles (mass, volume)
template <typename TF>
hipError_t launch_budget(hipStream_t st, Dims d, const int32_t* az, const double* h,
                         const TF* f, double* partial, int nblocks, double area, double* out);
// ntr tracers of one h (tracer t at f + t*tstride) in one pass: partial
// must hold 2*(ntr+1)*nblocks doubles, out ntr+1 (the masses, then the
// volume); each output equals launch_budget's with the same nblocks bit for
// bit
template <typename TF>
hipError_t launch_budget_batch(hipStream_t st, Dims d, const int32_t* az, int ntr, long tstride,
                               const double* h, const TF* f, double* partial, int nblocks,
                               double area, double* out);
// the final pass alone, for up to 8 domains in one launch: for domain i,
// out[q] = area * (sum of the npairs (hi, lo) pairs at partial + 2*q*npairs),
// q < nout, in k_budget_batch_final's order
struct BudgetPairs {
  const double* partial;
  double* out;
  long npairs;
  double area;
  int nout;
};
struct BudgetPairsSet {
  BudgetPairs d[8];
  int n;
};
hipError_t launch_budget_pairs_final(hipStream_t st, const BudgetPairsSet& S);

}  // namespace gt
