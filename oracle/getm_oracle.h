/*
 * getm_oracle.h -- CPU restatement of GETM's 3-D transport step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * code_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product never links or calls it.
 *
 * Every function restates one reference routine, with the same loop bounds,
 * the same masks and the same left-to-right operation order (build with
 * -ffp-contract=off), so that a faithful FP64 GPU kernel agrees bit for bit.
 * Citations are to getm-model/code.
 *
 * Array layout (include/cppdefs.h:146-159 of the reference): a 2-D field is
 * f(imin-2:imax+2, jmin-2:jmax+2), a 3-D field f(imin-2:imax+2,
 * jmin-2:jmax+2, 0:kmax), column-major with i fastest, FP64.  Masks are one
 * byte per cell (Fortran logical -> 0/1); az/au/av/ax are int32.
 */
#ifndef GETM_ORACLE_H
#define GETM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scheme / split / tag ids: src/2d/advection.F90:55-73,
 * src/3d/advection_3d.F90:42, src/futils/halo_zones.F90:24-28 */
enum { OR_NOADV = 0, OR_UPSTREAM = 1, OR_UPSTREAM_2DH = 2, OR_P2 = 3,
       OR_SUPERBEE = 4, OR_MUSCL = 5, OR_P2_PDM = 6, OR_J7 = 7, OR_FCT = 8,
       OR_P2_2DH = 9 };
enum { OR_NOSPLIT = 0, OR_FULLSPLIT = 1, OR_HALFSPLIT = 2, OR_HVSPLIT = 3 };
enum { OR_H_TAG = 10, OR_D_TAG = 20, OR_U_TAG = 31, OR_V_TAG = 32,
       OR_W_TAG = 33 };

/* error codes (the reference stops; the restatement returns) */
enum { OR_OK = 0, OR_EINVAL_TAG = 1, OR_EINVAL_SPLIT = 2,
       OR_EINVAL_SCHEME = 3, OR_ESCHEME_SPLIT = 4, OR_ENOMEM = 5 };

typedef struct {
    int imin, imax, jmin, jmax, kmax;
    double dx, dy, ard1;        /* ard1 = 1/(dx*dy), domain.F90:596 */
} or_dom;

/* t_adv_grid, src/2d/advection.F90:43-53 */
typedef struct {
    uint8_t *mask_uflux, *mask_vflux, *mask_uupdate, *mask_vupdate,
            *mask_finalise;
    int32_t *az;
} or_advgrid;

/* all three grids built by init_advection plus their storage */
typedef struct {
    or_advgrid H, U, V;
    uint8_t *store;
} or_advgrids;

/* a1: advection.F90:785-879 */
double or_adv_interfacial_reconstruction(int scheme, double cfl, double fuu,
                                         double fu, double fd);

/* a8: advection.F90:263-318; masks from domain.F90:294-346 */
int or_masks_from_az(const or_dom *d, const int32_t *az, int32_t *au,
                     int32_t *av, int32_t *ax);
int or_init_advection(const or_dom *d, const int32_t *az, const int32_t *au,
                      const int32_t *av, const int32_t *ax, or_advgrids *g);
void or_free_advection(or_advgrids *g);

/* a2/a3: adv_split_u.F90 / adv_split_v.F90 on one 2-D slice */
void or_adv_split_u(const or_dom *d, double dt, const double *f, double *fi,
                    double *Di, double *adv, const double *U,
                    const double *DU, double splitfac, int scheme, double AH,
                    const uint8_t *mask_flux, const uint8_t *mask_update);
void or_adv_split_v(const or_dom *d, double dt, const double *f, double *fi,
                    double *Di, double *adv, const double *V,
                    const double *DV, double splitfac, int scheme, double AH,
                    const uint8_t *mask_flux, const uint8_t *mask_update);
/* a4: adv_split_w.F90; returns the number of columns that wanted more than
 * itersmax iterations (the reference prints a warning per column) */
long or_adv_split_w(const or_dom *d, double dt, const double *f, double *fi,
                    double *hi, double *adv, const double *ww,
                    double splitfac, int scheme, int tag, const int32_t *az,
                    int itersmax);
/* a7: adv_upstream_2dh.F90 */
void or_adv_upstream_2dh(const or_dom *d, double dt, const double *f,
                         double *fi, double *Di, double *adv,
                         const double *U, const double *V, const double *Dn,
                         const double *DU, const double *DV, double AH,
                         const int32_t *az);
/* a6: advection.F90:362-702 (2-D driver, one slice).  Dires/advres may be
 * NULL.  *do_mirror is unused here (kept for the halo quirk). */
int or_do_advection(const or_dom *d, const or_advgrids *g, double dt,
                    double *f, const double *U, const double *V,
                    const double *DU, const double *DV, const double *Do,
                    const double *Dn, int split, int scheme, double AH,
                    int tag, double *Dires, double *advres);
/* a5: advection_3d.F90:120-598.  hires/advres may be NULL.  *nwarn receives
 * the count of columns that hit itersmax (may be NULL). */
int or_do_advection_3d(const or_dom *d, const or_advgrids *g, double dt,
                       double *f, const double *uu, const double *vv,
                       const double *ww, const double *hu, const double *hv,
                       const double *ho, const double *hn, int split,
                       int hscheme, int vscheme, double AH, int tag,
                       int adv_ver_iterations, double *hires, double *advres,
                       long *nwarn);

/* a12: eqstate.F90:127-149,252-261 (linear EOS over the full halo range) */
void or_eos_linear(const or_dom *d, const double *T, const double *S,
                   double *rho, double *buoy, double T0, double S0,
                   double dtr0, double dsr0);
/* full EOS, eqstate.F90:280-562: UNESCO 1980 one-atmosphere density,
 * Jackett et al. (2006) density at 0 and p dbar, and its derivatives wrt
 * salinity and potential temperature */
double or_rho_unesco80(double T, double S, int nonnegsalt);
void or_rho_from_theta(double s, double th, double p, double *dens0,
                       double *densp);
void or_eosall_from_theta(double s, double th, double p, double *rho_s,
                          double *rho_th);
/* do_eqstate, eqstate.F90:94-271 (method 1 linear, 2 UNESCO, 3 Jackett);
 * kmin = 1; alpha/beta may be NULL; returns 1 for a bad method */
int or_do_eqstate(const or_dom *d, const int32_t *az, int method,
                  int nonnegsalt, const double *T, const double *S,
                  const double *hn, double *rho, double *buoy, double *alpha,
                  double *beta, double T0, double S0, double dtr0, double dsr0);
/* a9: ip_blumberg_mellor.F90:8-152; zz is the module array, in/out */
void or_ip_blumberg_mellor(const or_dom *d, const int32_t *az,
                           const int32_t *au, const int32_t *av,
                           const double *H, const double *hn,
                           const double *hun, const double *hvn,
                           const double *buoy, double *zz, double *idpdx,
                           double *idpdy);
/* a10: ip_shchepetkin_mcwilliams.F90:8-291 */
void or_ip_shchepetkin_mcwilliams(const or_dom *d, const int32_t *az,
                                  const int32_t *au, const int32_t *av,
                                  const double *H, const double *hn,
                                  const double *hun, const double *hvn,
                                  const double *buoy, const double *sseo,
                                  double *zz, double *idpdx, double *idpdy);

/* a11: SUBSTR_INI_PRESS, internal_pressure.F90:235-258 (first: the SAVE'd
 * first-call flag, cleared on return) */
void or_substr_ini_press(const or_dom *d, int *first, double *idpdx,
                         double *idpdy, double *idpdx0, double *idpdy0);

/* a13: halo_zones.F90:81-187 serial semantics.  do_mirror_state carries the
 * SAVE'd flag of update_2d_halo (pass mirror = -1 for "absent"). */
void or_update_3d_halo(const or_dom *d, double *f1, const double *f2,
                       int mirror);
void or_update_2d_halo(const or_dom *d, double *f1, const double *f2,
                       int mirror, int *do_mirror_state);
/* TWOD exchange between tiles of one decomposed domain
 * (halo_mpi.F90:1095-1154): tiles[t] holds tile t of a px*py decomposition
 * (t = ty*px + tx), each with its own halo.  Each tile's ghost strips and
 * corners (HALO wide, k = 0..nk-1) are filled from its neighbours'
 * interiors; edges without a neighbour are left unchanged. */
void or_exchange_tiles(int px, int py, const int *ni, const int *nj, int nk,
                       double **tiles);
/* decomp_1d, halo_mpi.F90:566-591 */
void or_decomp_1d(int n, int np, int m, int *s, int *e);

/* a15: budgets.  sum over interior cells with az==1, k=1..kmax of
 * h*f*area (mass) and h*area (volume), error-free (double-double). */
void or_budget(const or_dom *d, const int32_t *az, const double *h,
               const double *f, double *mass, double *volume);

/* input producers (m3d.F90:446-548), Cartesian, kmin = 1 */
/* start_macro, start_macro.F90:54-86 */
void or_start_macro(const or_dom *d, int M, double min_depth, const double *H,
                    const double *HU, const double *HV, const double *z,
                    double *sseo, double *ssen, double *ssuo, double *ssun,
                    double *ssvo, double *ssvn, double *Dn, double *Dun,
                    double *Dvn, double *Uint, double *Vint);
/* the SAVE'd state of coordinates() (vert_cord 1 sigma, 2 z-level, 3
 * general; the namelist values ddu, ddl, d_gamma, gamma_surf, min_depth) */
void *or_coord_new(const or_dom *d, int vert_cord, double ddu, double ddl,
                   double d_gamma, int gamma_surf, double min_depth);
void or_coord_delete(void *state);
/* coordinates, coordinates.F90:66-107 with sigma_coordinates.F90:53-147 and
 * general_coordinates.F90:56-210; zc may be NULL; returns 0, 1 (general
 * coordinates not monotone) or 2 (vert_cord not 1, 2 or 3) */
int or_coordinates(const or_dom *d, void *state, double dt, double cord_relax,
                   double maxdepth, const double *H, const double *HU,
                   const double *HV, const double *sseo, const double *ssen,
                   const double *ssuo, const double *ssun, const double *ssvo,
                   const double *ssvn, double *ho, double *hn, double *huo,
                   double *hun, double *hvo, double *hvn, double *zc);
/* ww_momentum_3d, ww_momentum_3d.F90:66-119 */
void or_ww_momentum_3d(const or_dom *d, const int32_t *az, double dt,
                       const double *uu, const double *vv, const double *ho,
                       const double *hn, double *ww);

/* the step after advection: implicit vertical diffusion */
/* getm_tridiagonal, src/futils/tridiagonal.F90:8-51 */
void or_tridiagonal(int kmax, int fi, int lt, const double *au, const double *bu,
                    const double *cu, const double *du, double *value);
/* do_salinity's vertical diffusion, salinity.F90:346-393 */
void or_salinity_diffusion(const or_dom *d, const int32_t *az, double dt,
                           double cnpar, const double *hn, const double *nuh,
                           double *S);
/* do_temperature's solar radiation and vertical diffusion,
 * temperature.F90:463-551; rad is written */
void or_temperature_diffusion(const or_dom *d, const int32_t *az, double dt,
                              double cnpar, const double *hn, const double *nuh,
                              const double *swr, const double *shf,
                              const double *A, const double *g1, const double *g2,
                              const double *dry_z, const double *S,
                              double swr_bot_refl_frac, double swr_min_bot_frac,
                              double *rad, double *T);

/* a caller of the path: uv_advect_3d, src/3d/uv_advect_3d.F90:60-300
 * (J7 and the numerical analyses excluded) */
int or_uv_advect_3d(const or_dom *d, const or_advgrids *g, double dt,
                    const double *uu, const double *vv, const double *ww,
                    const double *ho, const double *hn, const double *hun,
                    const double *hvn, int split, int hscheme, int vscheme,
                    int adv_ver_iterations, double *uuEx, double *vvEx);

/* OpenMP thread count used by the restatement (0 = runtime default) */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
