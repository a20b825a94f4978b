/*
 * getm_transport.h -- C-ABI of the MI355X transport step (libgetm_transport.so).
 *
 * Drop-in boundary for GETM's 3-D baroclinic transport step
 * (getm-model/code).  Each entry point replaces one Fortran routine and keeps
 * its argument list and meaning; the reference routine is cited beside it.
 * A Fortran caller binds these through ISO_C_BINDING (the module in
 * code_amd/fortran/getm_transport_gpu.F90 keeps the reference's Fortran
 * signatures on top of them).
 *
 * Conventions
 *  - Fields use the reference layout f(imin-2:imax+2, jmin-2:jmax+2, 0:kmax),
 *    column-major, i fastest, FP64 (include/cppdefs.h:116-159).  A "slice"
 *    argument is one E2DFIELD level f(:,:,k).
 *  - Field pointers are DEVICE pointers (hipMalloc'd or gt_malloc'd); masks
 *    are passed once, as host arrays, to gt_create.
 *  - Every entry point returns 0 on success or a GT_E* code; the message of
 *    the last failure on the calling thread is gt_last_error().  The
 *    reference stops on the same conditions (advection_3d.F90:276,403,541,
 *    545,564; advection.F90:436,534,...).
 *  - Work is enqueued on the context's stream (gt_set_stream; default: the
 *    null stream).  Calls return once enqueued; gt_synchronize waits.
 *    gt_budget and gt_warnings synchronise because they return host values.
 *  - One context per domain; a context is not re-entrant (the reference's
 *    module state is not either) but distinct contexts are independent.
 */
#ifndef GETM_TRANSPORT_H
#define GETM_TRANSPORT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ids: advection.F90:55-73, advection_3d.F90:42, halo_zones.F90:24-28 */
enum { GT_NOADV = 0, GT_UPSTREAM = 1, GT_UPSTREAM_2DH = 2, GT_P2 = 3,
       GT_SUPERBEE = 4, GT_MUSCL = 5, GT_P2_PDM = 6, GT_J7 = 7, GT_FCT = 8,
       GT_P2_2DH = 9 };
enum { GT_NOSPLIT = 0, GT_FULLSPLIT = 1, GT_HALFSPLIT = 2, GT_HVSPLIT = 3 };
enum { GT_H_TAG = 10, GT_D_TAG = 20, GT_U_TAG = 31, GT_V_TAG = 32,
       GT_W_TAG = 33 };
/* the remaining halo-message tags of halo_zones.F90:25-27 (the ghost-fill
 * entry points accept any tag: one domain exchanges no messages) */
enum { GT_HU_TAG = 11, GT_HV_TAG = 12, GT_DU_TAG = 21, GT_DV_TAG = 22,
       GT_Z_TAG = 30 };
/* internal_pressure.F90:94-100 */
enum { GT_IP_BLUMBERG_MELLOR = 1, GT_IP_SHCHEPETKIN_MCWILLIAMS = 6 };

enum {
  GT_OK = 0,
  GT_EINVAL_TAG = 1,      /* 'do_advection_3d: tag is invalid' */
  GT_EINVAL_SPLIT = 2,    /* 'split is invalid' */
  GT_EINVAL_SCHEME = 3,   /* 'hscheme is invalid' / 'scheme is invalid' */
  GT_ESCHEME_SPLIT = 4,   /* 'hscheme not valid for split' */
  GT_ENOMEM = 5,
  GT_EHIP = 6,            /* a HIP runtime error (message has the detail) */
  GT_EARG = 7,            /* bad size, NULL pointer or context */
  GT_ENOTIMPL = 8         /* scheme outside this build (J7, FCT, P2_2DH) */
};

typedef struct gt_ctx gt_ctx;

const char *gt_last_error(void);
int gt_version(void);

/* ---- context: init_domain masks + init_advection + init_internal_pressure
 * Replaces the set-up of src/2d/advection.F90:204-354 (t_adv_grid masks),
 * src/3d/internal_pressure.F90:117-176 (zz allocation) on one domain.
 * az is required; au/av/ax may be NULL and are then derived from az as
 * src/domain/domain.F90:294-346 does.  Host arrays of E2DFIELD size. */
int gt_create(gt_ctx **ctx, int imin, int imax, int jmin, int jmax, int kmax,
              double dx, double dy, const int32_t *az, const int32_t *au,
              const int32_t *av, const int32_t *ax);
int gt_destroy(gt_ctx *ctx);
/* The context's masks, copied to host arrays of E2DFIELD size (any may be
 * NULL): az as given, au/av/ax as given or as derived from az at creation
 * (domain.F90:294-346, the module variables au, av, ax of domain.F90 that
 * the host model's other routines read).  Synchronous. */
int gt_get_masks(gt_ctx *ctx, int32_t *az, int32_t *au, int32_t *av,
                 int32_t *ax);
int gt_set_stream(gt_ctx *ctx, void *hip_stream);
int gt_synchronize(gt_ctx *ctx);
/* module variable adv_ver_iterations, advection_3d.F90:41 (default 1) */
int gt_set_adv_ver_iterations(gt_ctx *ctx, int iters);
/* columns that asked for more than itersmax vertical iterations since the
 * last call (the reference prints 'too many iterations needed',
 * adv_split_w.F90:219-231); synchronises */
long long gt_warnings(gt_ctx *ctx);

/* ---- advection --------------------------------------------------------- */
/* adv_interfacial_reconstruction, advection.F90:785-879, evaluated on n
 * device values (a check entry; the kernels inline the same function) */
int gt_adv_interfacial_reconstruction(gt_ctx *ctx, int scheme, long n,
                                      const double *cfl, const double *fuu,
                                      const double *fu, const double *fd,
                                      double *out);
/* adv_split_u, src/2d/adv_split_u.F90:8-295 (one slice; the mask pair of
 * the reference is chosen by tag: H/D/W -> adv_gridH, U -> adv_gridU,
 * V -> adv_gridV).  f and fi may be the same array. */
int gt_adv_split_u(gt_ctx *ctx, double dt, const double *f, double *fi,
                   double *Di, double *adv, const double *U, const double *DU,
                   double splitfac, int scheme, double AH, int tag);
/* adv_split_v, src/2d/adv_split_v.F90:8-145 */
int gt_adv_split_v(gt_ctx *ctx, double dt, const double *f, double *fi,
                   double *Di, double *adv, const double *V, const double *DV,
                   double splitfac, int scheme, double AH, int tag);
/* adv_split_w, src/3d/adv_split_w.F90:7-360 (full I3DFIELD arrays) */
int gt_adv_split_w(gt_ctx *ctx, double dt, const double *f, double *fi,
                   double *hi, double *adv, const double *ww,
                   double splitfac, int scheme, int tag, int itersmax);
/* adv_upstream_2dh, src/2d/adv_upstream_2dh.F90:8-169 (one slice) */
int gt_adv_upstream_2dh(gt_ctx *ctx, double dt, const double *f, double *fi,
                        double *Di, double *adv, const double *U,
                        const double *V, const double *Dn, const double *DU,
                        const double *DV, double AH, int tag);
/* The four routines above with the masks the reference's explicit
 * interfaces take as arguments instead of a tag (advection.F90:91-194,
 * advection_3d.F90:55-68): the caller's t_adv_grid masks as device bytes
 * (1 = .true.) over E2DFIELD, as seen through the reference's pointer
 * views, and adv_grid%az as device int32.  adv_split_v reads rows 0..NY-2
 * of mask_flux only (the reference declares it with NY-1 rows). */
int gt_adv_split_u_m(gt_ctx *ctx, double dt, const double *f, double *fi,
                     double *Di, double *adv, const double *U, const double *DU,
                     double splitfac, int scheme, double AH,
                     const uint8_t *mask_flux, const uint8_t *mask_update);
int gt_adv_split_v_m(gt_ctx *ctx, double dt, const double *f, double *fi,
                     double *Di, double *adv, const double *V, const double *DV,
                     double splitfac, int scheme, double AH,
                     const uint8_t *mask_flux, const uint8_t *mask_update);
int gt_adv_split_w_m(gt_ctx *ctx, double dt, const double *f, double *fi,
                     double *hi, double *adv, const double *ww,
                     double splitfac, int scheme, int tag, const int32_t *az,
                     int itersmax);
int gt_adv_upstream_2dh_m(gt_ctx *ctx, double dt, const double *f, double *fi,
                          double *Di, double *adv, const double *U,
                          const double *V, const double *Dn, const double *DU,
                          const double *DV, double AH, const int32_t *az);
/* do_advection, src/2d/advection.F90:362-702 (one slice).  Dires and advres
 * are the optional outputs (NULL = absent). */
int gt_do_advection(gt_ctx *ctx, double dt, double *f, const double *U,
                    const double *V, const double *DU, const double *DV,
                    const double *Do, const double *Dn, int split, int scheme,
                    double AH, int tag, double *Dires, double *advres);
/* do_advection_3d, src/3d/advection_3d.F90:120-598.  hires and advres are
 * the optional outputs (NULL = absent). */
int gt_do_advection_3d(gt_ctx *ctx, double dt, double *f, const double *uu,
                       const double *vv, const double *ww, const double *hu,
                       const double *hv, const double *ho, const double *hn,
                       int split, int hscheme, int vscheme, double AH, int tag,
                       double *hires, double *advres);

/* ---- input producers: what integrate_3d runs before the tracers
 * (m3d.F90:446-548).  Cartesian metrics (dx, dy of the context); kmin = 1.
 * 2-D fields are E2DFIELD, 3-D fields I3DFIELD, device pointers. */
/* start_macro, src/3d/start_macro.F90:54-86: sseo = ssen, ssen = z,
 * Dn = ssen + H; the U/V-point elevations are the means of the old and new
 * elevations clipped at -HU + min_depth (-HV + min_depth), ssuo/ssvo the
 * previous ones, Dun/Dvn the depths; Uint and Vint are divided by M. */
int gt_start_macro(gt_ctx *ctx, int M, double min_depth, const double *H,
                   const double *HU, const double *HV, const double *z,
                   double *sseo, double *ssen, double *ssuo, double *ssun,
                   double *ssvo, double *ssvn, double *Dn, double *Dun,
                   double *Dvn, double *Uint, double *Vint);
/* the domain-namelist values coordinates() reads (domain.F90: vert_cord,
 * ddu, ddl, d_gamma, gamma_surf, min_depth).  vert_cord 1 (sigma), 2
 * (z-levels: zc only, as the reference, which codes nothing else for them)
 * or 3 (general); 4 returns GT_ENOTIMPL with the reference's 'hybrid
 * coordinates not coded yet', 5 (adaptive) is outside this build.  Resets
 * the first-call state. */
int gt_set_coordinates(gt_ctx *ctx, int vert_cord, double ddu, double ddl,
                       double d_gamma, int gamma_surf, double min_depth);
/* coordinates, src/3d/coordinates.F90:66-107 with
 * sigma_coordinates.F90:53-147 and general_coordinates.F90:56-210.  The
 * first call initialises (level arrays; for general coordinates the gamma
 * levels of every column and the initial hn, huo = hun, hvo = hvn), every
 * call updates ho, hn, huo, hun, hvo, hvn; zc (may be NULL) gets the layer
 * centres.  Folded gamma levels return GT_EARG (the reference stops,
 * general_coordinates.F90:92-98); the first call synchronises. */
int gt_coordinates(gt_ctx *ctx, double dt, double cord_relax, double maxdepth,
                   const double *H, const double *HU, const double *HV,
                   const double *sseo, const double *ssen, const double *ssuo,
                   const double *ssun, const double *ssvo, const double *ssvn,
                   double *ho, double *hn, double *huo, double *hun,
                   double *hvo, double *hvn, double *zc);
/* ww_momentum_3d, src/3d/ww_momentum_3d.F90:66-119: ww(k) for k =
 * 1..kmax-1 on interior water columns, upward from ww(0) (read); the serial
 * 3-D halo update after it is a no-op */
int gt_ww_momentum_3d(gt_ctx *ctx, double dt, const double *uu,
                      const double *vv, const double *ho, const double *hn,
                      double *ww);

/* ---- a caller of the path: momentum advection ------------------------------
 * uv_advect_3d, src/3d/uv_advect_3d.F90:8-466, for every scheme but J7
 * (GT_ENOTIMPL) and without the numerical analyses: u = uu/hun on the U
 * control volumes, with the transports, ww and the layer heights averaged
 * onto them, advected by do_advection_3d (tag U_TAG, AH = 0), advection term
 * into uuEx; the same on the V volumes into vvEx.  The context keeps six
 * work fields (the reference's automatic arrays). */
int gt_uv_advect_3d(gt_ctx *ctx, double dt, const double *uu, const double *vv,
                    const double *ww, const double *ho, const double *hn,
                    const double *hun, const double *hvn, int split,
                    int hscheme, int vscheme, double *uuEx, double *vvEx);

/* ---- the step after advection: implicit vertical diffusion ---------------
 * The part of do_salinity / do_temperature that follows their
 * do_advection_3d call, on interior water columns (az == 1), kmax > 1:
 * Crank-Nicolson (cnpar) diffusion with nuh + the molecular diffusivity
 * (parameters.F90:26-27: avmolt = 1.4e-7, avmols = 1.1e-9, FP32 literals),
 * solved per column with getm_tridiagonal. */
/* salinity.F90:346-393 */
int gt_salinity_diffusion(gt_ctx *ctx, double dt, double cnpar,
                          const double *hn, const double *nuh, double *S);
/* temperature.F90:463-551: rad = swr (A exp(-z/g1) + (1-A) exp(-z/g2)) on
 * columns with az >= 1 (written, k = 0..kmax), the sea-ice rule (no cooling
 * flux where T(kmax) <= -0.0575 S(kmax)), the bottom reflection
 * (swr_bot_refl_frac > 0 and rad(0)/rad(kmax) > swr_min_bot_frac), and the
 * diffusion with dry_z dt (rad divergence + shf at the surface)/(rho_0 cp) as
 * source.  swr, shf, A, g1, g2, dry_z are 2-D; S is read at kmax. */
int gt_temperature_diffusion(gt_ctx *ctx, double dt, double cnpar,
                             const double *hn, const double *nuh,
                             const double *swr, const double *shf,
                             const double *A, const double *g1,
                             const double *g2, const double *dry_z,
                             const double *S, double swr_bot_refl_frac,
                             double swr_min_bot_frac, double *rad, double *T);
/* getm_tridiagonal, src/futils/tridiagonal.F90:8-51, over n independent
 * systems, rows fi..lt (0 <= fi < lt): au sub-, bu main, cu super-diagonal,
 * du right-hand side, value the solution; element k of system s at
 * [k*n + s] (level-major, so the systems run across a wavefront) */
int gt_tridiagonal_batch(gt_ctx *ctx, long n, int fi, int lt,
                         const double *au, const double *bu, const double *cu,
                         const double *du, double *value);

/* ---- restart fields: the NetCDF hotstart file ---------------------------
 * create_restart_ncdf.F90 / write_restart_ncdf.F90 / read_restart_ncdf.F90
 * for the fields a caller names (e.g. z, hn, T, S, uu, vv, ww, ssen, sseo):
 * NetCDF classic format written and read natively, dimensions xax, yax,
 * zax = kmax+1, the scalars loop, julianday, secondsofday (int) and timestep
 * (double) of create_restart_ncdf.F90:94-104, FP64 variables over (x, y) or
 * (x, y, z) in the reference's Fortran order, the title attribute.  Writing:
 * with_halos = 0 stores imin..imax x jmin..jmax (the default build), 1 the
 * whole array (_WRITE_HOT_HALOS_); time may be NULL (zeros).  64-bit
 * offsets when the file exceeds 2 GiB.  Reading requires the four scalars
 * (open_restart_ncdf.F90:48-62) and returns them in time (may be NULL); it
 * fills the interior cells only, from an interior-sized file or, with
 * with_halos = 1 (__READ_HOT_HALOS_, read_restart_ncdf.F90:83-101), also
 * from a halo-inclusive one; halo cells of the arrays are left for the
 * caller's update_2d/3d_halo.  The reference reader also needs 'z'. */
typedef struct {
  const char *name; /* NetCDF variable name, as the reference's */
  int ndim;         /* 2: E2DFIELD, 3: I3DFIELD */
  double *data;     /* the whole array, halos included */
} gt_restart_field;
typedef struct {
  int loop, julianday, secondsofday;
  double timestep;
} gt_restart_time;
/* data are host arrays */
int gt_restart_write_host(const char *path, int imin, int imax, int jmin,
                          int jmax, int kmax, int with_halos,
                          const gt_restart_time *time, int n,
                          const gt_restart_field *fields);
int gt_restart_read_host(const char *path, int imin, int imax, int jmin,
                         int jmax, int kmax, int with_halos,
                         gt_restart_time *time, int n,
                         const gt_restart_field *fields);
/* data are device arrays of the context's domain (synchronises) */
int gt_write_restart(gt_ctx *ctx, const char *path, int with_halos,
                     const gt_restart_time *time, int n,
                     const gt_restart_field *fields);
int gt_read_restart(gt_ctx *ctx, const char *path, int with_halos,
                    gt_restart_time *time, int n,
                    const gt_restart_field *fields);

/* ---- internal pressure ------------------------------------------------- */
/* do_eqstate with eqstate_method=1 (linear), src/3d/eqstate.F90:127-149 and
 * the buoyancy of :252-261.  rho may be NULL.  The reference defaults are
 * T0=10, S0=33.75, dtr0=-0.17, dsr0=0.78 as FP32 literals. */
int gt_eos_linear(gt_ctx *ctx, const double *T, const double *S, double *rho,
                  double *buoy, double T0, double S0, double dtr0,
                  double dsr0);
/* do_eqstate, src/3d/eqstate.F90:94-271 (the routine is argument-less; its
 * module variables are arguments here).  method = eqstate_method: 1 linear
 * (rho on every cell, alpha = dtr0 and beta = dsr0 on the whole array), 2
 * UNESCO 1980, 3 Jackett et al. (2006) (rho on water columns az > 0; land
 * keeps its rho).  For 2 and 3, alpha/beta are d rho/d theta and d rho/d S at
 * the layer interfaces k = 1..kmax-1 (zero at kmax) from the interface means
 * of T and S at the pressure p = sum of hn above (kmin = 1, as sigma and
 * general coordinates set it).  buoy = -g/rho_0 (rho - rho_0) on every cell.
 * alpha and beta may both be NULL (the _OLD_BVF_ build); hn is read only for
 * them.  nonnegsalt = 1 is the NONNEGSALT build (negative S taken as 0 in
 * rho).  A bad method returns GT_EARG ('init_eqstate(): not a valid
 * eqstate_method', eqstate.F90:75-77). */
int gt_do_eqstate(gt_ctx *ctx, int method, int nonnegsalt, const double *T,
                  const double *S, const double *hn, double *rho,
                  double *buoy, double *alpha, double *beta, double T0,
                  double S0, double dtr0, double dsr0);
/* the module's point functions, elementwise over n device points:
 * rho_from_theta_unesco80 (eqstate.F90:280-329), rho_from_theta (:338-423;
 * p may be NULL for p = 0; densp may be NULL), eosall_from_theta (:432-562) */
int gt_rho_from_theta_unesco80(gt_ctx *ctx, long n, const double *T,
                               const double *S, double *rho, int nonnegsalt);
int gt_rho_from_theta(gt_ctx *ctx, long n, const double *s, const double *th,
                      const double *p, double *dens0, double *densp);
int gt_eosall_from_theta(gt_ctx *ctx, long n, const double *s,
                         const double *th, const double *p, double *rho_s,
                         double *rho_th);
/* ip_blumberg_mellor, src/3d/ip_blumberg_mellor.F90:8-152.  The module
 * variables it reads are arguments here; zz is the module array
 * (internal_pressure.F90:66-72), written where the reference writes it. */
int gt_ip_blumberg_mellor(gt_ctx *ctx, const double *H, const double *hn,
                          const double *hun, const double *hvn,
                          const double *buoy, double *zz, double *idpdx,
                          double *idpdy);
/* ip_shchepetkin_mcwilliams, src/3d/ip_shchepetkin_mcwilliams.F90:8-291 */
int gt_ip_shchepetkin_mcwilliams(gt_ctx *ctx, const double *H,
                                 const double *hn, const double *hun,
                                 const double *hvn, const double *buoy,
                                 const double *sseo, double *zz,
                                 double *idpdx, double *idpdy);
/* do_eqstate with the linear EOS (eqstate.F90:127-149,252-261) followed by
 * ip_blumberg_mellor (methods & 1) and/or ip_shchepetkin_mcwilliams
 * (methods & 2), fused: the buoyancy is formed from T and S where the
 * pressure gradients read it and is not stored (a caller that needs buoy
 * calls gt_eos_linear), the two methods' heights come from one read of hn,
 * and every gradient is bit for bit that of gt_eos_linear followed by
 * gt_ip_blumberg_mellor / gt_ip_shchepetkin_mcwilliams on the same zz
 * arrays.  The outputs of a method not asked for may be NULL; sseo is read
 * by Shchepetkin-McWilliams only.  No counterpart in the reference as one
 * routine (do_eqstate at m3d.F90:607 feeds the next step's
 * do_internal_pressure at m3d.F90:532). */
int gt_eos_internal_pressure(gt_ctx *ctx, int methods, const double *T,
                             const double *S, double T0, double S0,
                             double dtr0, double dsr0, const double *H,
                             const double *hn, const double *hun,
                             const double *hvn, const double *sseo,
                             double *zz_bm, double *idpdx_bm, double *idpdy_bm,
                             double *zz_sm, double *idpdx_sm, double *idpdy_sm);
/* the same for any eqstate_method (1 linear, 2 UNESCO 1980, 3 Jackett et
 * al.; eqstate.F90:94-261, nonnegsalt as in gt_do_eqstate): every gradient
 * is bit for bit that of gt_do_eqstate followed by gt_ip_blumberg_mellor /
 * gt_ip_shchepetkin_mcwilliams.  Methods 2 and 3 read rho on land columns
 * (az = 0) only: do_eqstate leaves rho there and forms the buoyancy from it
 * (eqstate.F90:154,200,252-261), and Shchepetkin-McWilliams' harmonic means
 * reach one column past a water face; rho may be NULL for method 1, and is
 * never written (a caller that needs rho or buoy calls gt_do_eqstate).
 * T0, S0, dtr0, dsr0 are read by method 1 only.  gt_eos_internal_pressure
 * is this call with eqstate_method 1. */
int gt_eqstate_internal_pressure(gt_ctx *ctx, int eqstate_method,
                                 int nonnegsalt, int methods, const double *T,
                                 const double *S, const double *rho, double T0,
                                 double S0, double dtr0, double dsr0,
                                 const double *H, const double *hn,
                                 const double *hun, const double *hvn,
                                 const double *sseo, double *zz_bm,
                                 double *idpdx_bm, double *idpdy_bm,
                                 double *zz_sm, double *idpdx_sm,
                                 double *idpdy_sm);
/* do_internal_pressure, src/3d/internal_pressure.F90:181-274, with the
 * context's own zz (zero at creation, as init_internal_pressure leaves it).
 * sseo is only read by method 6. */
int gt_do_internal_pressure(gt_ctx *ctx, int ip_method, const double *H,
                            const double *hn, const double *hun,
                            const double *hvn, const double *buoy,
                            const double *sseo, double *idpdx, double *idpdy);
/* the SUBSTR_INI_PRESS build of do_internal_pressure (internal_pressure.F90:
 * 132-141,235-258), as a context switch instead of a compile option: with
 * on = 1 the next gt_do_internal_pressure keeps its gradients (interior,
 * all levels) and returns zero there; later calls subtract those first
 * gradients.  Switching on allocates the two stored fields and restarts the
 * first-call state; 0 switches it off. */
int gt_set_substr_ini_press(gt_ctx *ctx, int on);

/* ---- ghost cells --------------------------------------------------------- */
/* update_3d_halo, src/futils/halo_zones.F90:141-187, serial semantics:
 * mirror = 1 copies the 1-wide edges, 0 or -1 (absent) does nothing. */
int gt_update_3d_halo(gt_ctx *ctx, double *f1, const double *f2, int tag,
                      int mirror);
/* update_2d_halo, halo_zones.F90:81-132; mirror = -1 means "absent" and
 * uses the flag the previous call left (the reference's SAVE'd do_mirror) */
int gt_update_2d_halo(gt_ctx *ctx, double *f1, const double *f2, int tag,
                      int mirror);
/* wait_halo, halo_zones.F90:196-217: waits for the context's stream */
int gt_wait_halo(gt_ctx *ctx, int tag);
/* The exchange of halo_mpi.F90:1095-1154 (TWOD, HALO-wide strips and
 * corners, levels 0..nk-1) between the px*py tiles of one decomposed domain
 * held in one device memory.  tiles[t] (device pointer, t = ty*px+tx) has
 * interior ni[t] x nj[t]; tiles in one row share nj, in one column share ni.
 * Edges without a neighbour are left unchanged (MPI_PROC_NULL).  At most 64
 * tiles.  Enqueued on the given stream (NULL = null stream); allocates
 * nothing and returns without waiting. */
int gt_exchange_tiles(int px, int py, const int *ni, const int *nj, int nk,
                      double *const *tiles, void *hip_stream);

/* ---- batched multi-tracer step (no reference counterpart) ---------------
 * The HALFSPLIT step of do_advection_3d (advection_3d.F90:439-547) for
 * ntracers tracers that share one velocity field, over ndom domains (e.g.
 * uneven subdomains), in one fused launch per 8 domains.  Tracer t of a
 * domain is the I3DFIELD at f_in + t*(kmax+1)*NY*NX; the result goes to
 * f_out (which must not overlap f_in).  The layer heights stay FP64; hires
 * (optional) receives the end of the layer-height trajectory.
 * _f32: tracer and flux arithmetic in FP32, the update (Dio*f-dt*advn)/Di in
 * FP64.  _f64: the reference arithmetic, bit-identical to do_advection_3d.
 * Vertical sub-cycling is not applied (adv_ver_iterations = 1).  Each
 * domain needs its own context (the kernels' work fields live in it; a
 * context given twice is rejected), and those work fields are shared with
 * the context's own stream: order the two streams when both are used. */
typedef struct {
  gt_ctx *ctx;
  const void *f_in;
  void *f_out;
  const double *uu, *vv, *ww, *hu, *hv, *ho;
  double *hires;
  /* optional (device, ntracers + 1 doubles, NULL = none): the FP64 budgets
   * of the result formed inside the step, tracer t's mass
   * sum(hires*f_out*dx*dy) over the interior water cells (az == 1) of levels
   * 1..kmax, then the volume sum(hires*dx*dy); each lane sums its column's
   * products over its strip's rows in FP64, a wavefront adds its lanes in
   * an FP64 tree and the strip-level partials are added as double-doubles,
   * so for terms of one sign the result is within (rows + 6) roundings of
   * the error-free gt_budget_batch_f32.  Tags H, D, W (the T points) only. */
  double *budget;
} gt_batch_domain;
int gt_adv3d_batch_f32(int ndom, const gt_batch_domain *doms, int ntracers,
                       double dt, int hscheme, int vscheme, double AH, int tag,
                       void *hip_stream);
int gt_adv3d_batch_f64(int ndom, const gt_batch_domain *doms, int ntracers,
                       double dt, int hscheme, int vscheme, double AH, int tag,
                       void *hip_stream);
/* kernel choice: the two HALFSPLIT strip kernels (fused=1) or the
 * per-sub-step kernels (0; gt_adv3d_batch_f32 then runs the FP32 sub-steps
 * one kernel each, bit-identical to the strip kernels, and accepts
 * hscheme != vscheme; gt_adv3d_batch_f64 always runs the strip kernels);
 * rows per strip (default: 16, or 8 on grids where 16-row strips give fewer
 * than 4096 strip-level items): tile 0, 1, 2 give 16, 32, 8 rows
 * to both kernels, tile 3..11 gives {16, 32, 8}[(tile-3)%3] rows to the first
 * (u/2 v/2) and {16, 32, 8}[(tile-3)/3] to the second (w v/2 u/2); tracers
 * per wavefront (group: FP32 1, 3, 6 or 9; FP64 1, or 2 for any other value
 * when there are two or more tracers) */
int gt_set_tuning(gt_ctx *ctx, int fused, int tile, int group);
/* reverse mode (gt_adv3d_adjoint*): keep_states 1 keeps the four sub-step
 * states of every step of a checkpoint segment while the segment is
 * recomputed, so the reverse steps do not recompute them (work memory
 * 4*c tracer sets instead of 4, and, with per-step inputs, c sets of layer
 * heights instead of one, c = min(ckpt_every, nsteps)); 0 recomputes them per
 * reverse step; -1 (default) keeps them when that needs at most half of the
 * free device memory.  Results are the same bit for bit.  No counterpart in
 * the reference (its reverse mode is this build's extension). */
int gt_set_adjoint_mode(gt_ctx *ctx, int keep_states);
/* sizes the context's work fields for the transport step up front (the
 * HALFSPLIT step with advres, batched steps of up to ntracers_f32 FP32 or
 * ntracers_f64 FP64 tracers with their budgets formed in the step,
 * NOSPLIT/FULLSPLIT/HVSPLIT buffers, SMcW's P, the budget partials, and the
 * sub-cycling columns if adv_ver_iterations >
 * 1), so later calls of those entry points neither allocate nor synchronise
 * the device to grow a field.  Without it they allocate on first use.  The
 * work memory of the transport step replaces the reference's automatic
 * arrays (do_advection_3d fi/hi/adv, SMcW dR/dZ/P). */
int gt_reserve(gt_ctx *ctx, int ntracers_f32, int ntracers_f64);
/* autotuner: times gt_adv3d_batch_f32 (f32=1) or _f64 over the nine strip
 * height pairs (tile 3..11) x tracers per wavefront {1, 3, 6, 9} (FP64:
 * {1, 2}), reps timed launches each, in two passes over the candidates,
 * and keeps the fastest on every context.  table (optional, 36 floats)
 * receives each pair's ms (the lower of its two passes' means) in sweep
 * order (tile outer).  f_out is overwritten, f_in is not touched. */
int gt_autotune_batch(int ndom, const gt_batch_domain *doms, int ntracers,
                      double dt, int hscheme, int vscheme, double AH, int tag,
                      int f32, int reps, void *hip_stream, int *best_tile,
                      int *best_group, float *best_ms, float *table);

/* ---- reverse mode (no reference counterpart) -----------------------------
 * Adjoint of nsteps HALFSPLIT steps of do_advection_3d (advection_3d.F90:
 * 439-547, adv_ver_iterations = 1) with respect to the tracer, velocities
 * and layer heights held fixed.  Runs the forward from f0 (result in f_final,
 * optional), keeping a checkpoint every ckpt_every steps, then sweeps back:
 * lam0 = (d f_final / d f0)^T lam_final.  The limiter enters through the
 * derivative of the branch the forward took.  The reverse sweep recomputes
 * each segment from its checkpoint and applies the transpose of each
 * sub-step; the states before the sub-steps come from the per-sub-step
 * kernels (bit-identical to the strip kernels), kept for the whole segment
 * or recomputed per reverse step (gt_set_adjoint_mode).  Work memory in FP64
 * fields, with c = min(ckpt_every, nsteps): 5 + 2 + ceil(nsteps/ckpt_every)
 * + c - 1, plus 4 (states recomputed) or 4*c (states kept). */
int gt_adv3d_adjoint(gt_ctx *ctx, int nsteps, int ckpt_every, double dt,
                     const double *f0, double *f_final, const double *lam_final,
                     double *lam0, const double *uu, const double *vv,
                     const double *ww, const double *hu, const double *hv,
                     const double *ho, int hscheme, int vscheme, double AH,
                     int tag);
/* the same for ntracers FP32 tracers sharing the velocities, in the
 * arithmetic of gt_adv3d_batch_f32 (tracer t of f0, f_final, lam_final and
 * lam0 at + t*(kmax+1)*NY*NX; layer heights FP64; adv_ver_iterations = 1).
 * The forward sweep runs the strip kernels (fused = 1 and hscheme == vscheme)
 * or the per-sub-step kernels, which give the same states bit for bit; the
 * reverse sweep recomputes each step's sub-step states from the checkpoint
 * (kept per segment or per reverse step, gt_set_adjoint_mode) and applies
 * the transpose of each FP32 sub-step, the limiter through the derivative of
 * the branch term the FP32 forward selected.  Work memory, with c =
 * min(ckpt_every, nsteps): 5 FP64 fields + (2 + ceil(nsteps/ckpt_every) +
 * c - 1 + 4 or 4*c) tracer sets of ntracers FP32 fields. */
int gt_adv3d_adjoint_f32(gt_ctx *ctx, int ntracers, int nsteps, int ckpt_every,
                         double dt, const float *f0, float *f_final,
                         const float *lam_final, float *lam0, const double *uu,
                         const double *vv, const double *ww, const double *hu,
                         const double *hv, const double *ho, int hscheme,
                         int vscheme, double AH, int tag);
/* the two above with inputs that change from step to step (a tide, the
 * producers of each macro step): uu[t], vv[t], ww[t], hu[t], hv[t], ho[t]
 * (host arrays of nsteps device pointers) are the inputs of step t; the layer
 * heights of the sub-steps are formed per step (with the states kept, once
 * per step of a segment: min(ckpt_every, nsteps) sets of 5 FP64 fields).  With the same
 * pointers for every step the result equals gt_adv3d_adjoint / _f32 bit for
 * bit. */
int gt_adv3d_adjoint_seq(gt_ctx *ctx, int nsteps, int ckpt_every, double dt,
                         const double *f0, double *f_final,
                         const double *lam_final, double *lam0,
                         const double *const *uu, const double *const *vv,
                         const double *const *ww, const double *const *hu,
                         const double *const *hv, const double *const *ho,
                         int hscheme, int vscheme, double AH, int tag);
int gt_adv3d_adjoint_seq_f32(gt_ctx *ctx, int ntracers, int nsteps,
                             int ckpt_every, double dt, const float *f0,
                             float *f_final, const float *lam_final,
                             float *lam0, const double *const *uu,
                             const double *const *vv,
                             const double *const *ww,
                             const double *const *hu,
                             const double *const *hv,
                             const double *const *ho, int hscheme,
                             int vscheme, double AH, int tag);
/* transpose of ip_blumberg_mellor (ip_blumberg_mellor.F90:99-149) with
 * respect to buoy, zz/hun/hvn from the forward call: buoy_bar is overwritten
 * (kmax <= 64) */
int gt_ip_blumberg_mellor_adjoint(gt_ctx *ctx, const double *hun,
                                  const double *hvn, const double *zz,
                                  const double *idpdx_bar,
                                  const double *idpdy_bar, double *buoy_bar);
/* transpose of ip_shchepetkin_mcwilliams (ip_shchepetkin_mcwilliams.F90:
 * 90-286) with respect to buoy at the forward point: hn, hun, hvn, sseo and
 * the buoyancy of the forward call, zz as that call wrote it; the harmonic
 * means enter through their derivatives (zero where the forward's eps test
 * failed).  buoy_bar is overwritten (kmax <= 64). */
int gt_ip_shchepetkin_mcwilliams_adjoint(gt_ctx *ctx, const double *hn,
                                         const double *hun, const double *hvn,
                                         const double *buoy, const double *sseo,
                                         const double *zz,
                                         const double *idpdx_bar,
                                         const double *idpdy_bar,
                                         double *buoy_bar);
/* transpose of the linear EOS + buoyancy (eqstate.F90:127-149,252-261):
 * T_bar and S_bar (either may be NULL) are overwritten on levels 1..kmax */
int gt_eos_linear_adjoint(gt_ctx *ctx, const double *buoy_bar, double *T_bar,
                          double *S_bar, double dtr0, double dsr0);

/* ---- budgets (no reference counterpart) ---------------------------------
 * mass = sum over interior water cells (az==1), k=1..kmax of h*f*dx*dy and
 * volume = sum of h*dx*dy, accumulated error-free in double-double.  f may be
 * NULL (volume only; mass then equals volume).  Synchronises. */
int gt_budget(gt_ctx *ctx, const double *h, const double *f, double *mass,
              double *volume);
/* the same with an FP32 tracer (each h*f product formed exactly in FP64) */
int gt_budget_f32(gt_ctx *ctx, const double *h, const float *f, double *mass,
                  double *volume);
/* the budgets of ntracers FP32 tracers sharing h (tracer t at
 * f + t*(kmax+1)*NY*NX, as gt_adv3d_batch_f32 lays them out) in one pass
 * that reads h once per six tracers: out (device memory, ntracers + 1
 * doubles) receives the masses, then the volume; each equals gt_budget_f32's
 * bit for bit.  Asynchronous on the context's stream; the partial sums live
 * in the context (grown on first use, or sized by gt_reserve). */
int gt_budget_batch_f32(gt_ctx *ctx, int ntracers, const double *h,
                        const float *f, double *out);
/* gt_budget_batch_f32 that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  No counterpart in the reference.  The partial
 * sums go to a static device array of 2*(32+1)*1024 doubles (528 KiB) that
 * each GPU holds until the process ends, not to the context: at most 32
 * tracers (more return GT_EARG).  Same kernels, same results bit for bit;
 * asynchronous on the context's stream, no synchronisation.  All calls on one
 * GPU share its array: calls that overlap on one GPU (two streams, two
 * threads) race. */
int gt_budget_batch_f32_noalloc(gt_ctx *ctx, int ntracers, const double *h,
                                const float *f, double *out);

/* ---- device memory helpers for callers without a HIP binding ---------- */
int gt_malloc(void **ptr, size_t bytes);
int gt_free(void *ptr);
int gt_memcpy_h2d(void *dst, const void *src, size_t bytes);
int gt_memcpy_d2h(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif
