/*
 * getm_oracle.c -- CPU restatement of GETM's 3-D transport step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, and the CPU
 * baseline that bench.py times beside it.  Never linked into the product.
 *
 * Each routine follows one reference routine of getm-model/code (cited per
 * function) with its loop bounds, masks and left-to-right evaluation order.
 * Build with -O2 -ffp-contract=off so no multiply-add is fused.  OpenMP
 * splits the outer j loops as the reference's !$OMP DO directives do; no
 * result depends on the thread count (there are no reductions except in
 * or_budget, which is serial and error-free).
 *
 * Every function is pinned by tests/test_oracle_*.py (closed forms,
 * published values, invariants, reflection symmetry, the reference's own
 * check values): the Shchepetkin-McWilliams brackets against exact
 * cubic-Hermite integrals, the W-point column, the sub-cycling restart,
 * and the NOSPLIT/HVSPLIT/FULLSPLIT and 2-D driver compositions against
 * textbook operators (tests/test_oracle_pins_branches.py).  What no check
 * fixes beyond this restatement's reading of the source: a halo-row loop
 * bound of adv_split_u/v wrong by the same amount on both edges
 * (reflection symmetry catches every asymmetric bound error).
 */
#include "getm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define HALO 2

/* ---- layout helpers -------------------------------------------------- */
#define DOMVARS(d)                                                          \
    const int il = (d)->imin, ih = (d)->imax, jl = (d)->jmin,               \
              jh = (d)->jmax, kmax = (d)->kmax;                             \
    const long NX = (long)(ih - il + 1 + 2 * HALO);                         \
    const long NY = (long)(jh - jl + 1 + 2 * HALO);                         \
    const long NXY = NX * NY;                                               \
    (void)kmax; (void)NXY
/* Fortran (i,j[,k]) -> linear offset */
#define IJ(i, j) ((long)((i) - (il - HALO)) + NX * (long)((j) - (jl - HALO)))
#define IJK(i, j, k) (IJ(i, j) + NXY * (long)(k))

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_NT num_threads(or_get_threads())

/* scratch buffers that stand in for the reference's automatic arrays, kept
 * between calls so the timed baseline does not page-fault fresh memory */
enum { SCR_FLUX = 0, SCR_VFLUX, SCR_CMIN, SCR_CMAX, SCR_FI, SCR_HI, SCR_ADV,
       SCR_DR, SCR_DZ, SCR_P, SCR_FC, SCR_DZX, SCR_DRX, SCR_N };
static double *g_scr[SCR_N];
static size_t g_scr_n[SCR_N];
static double *scr(int slot, size_t n)
{
    if (g_scr_n[slot] < n) {
        free(g_scr[slot]);
        g_scr[slot] = (double *)malloc(n * sizeof(double));
        g_scr_n[slot] = g_scr[slot] ? n : 0;
    }
    return g_scr[slot];
}

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }

/* ---- a1: limiter ------------------------------------------------------ */
/* advection.F90:854-879 */
static double reconstruction_p2(double cfl, double fu, double deltafu,
                                double deltaf)
{
    const double one6th = 1.0 / 6;
    double x = one6th * (1.0 - 2.0 * cfl);
    return fu + 0.5 * (1.0 - cfl) * ((0.5 + x) * deltaf + (0.5 - x) * deltafu);
}

/* advection.F90:785-845 */
double or_adv_interfacial_reconstruction(int scheme, double cfl, double fuu,
                                         double fu, double fd)
{
    const double one6th = 1.0 / 6;
    double deltaf = fd - fu;
    double deltafu = fu - fuu;
    double limiter = 0.0;
    if (deltaf * deltafu > 0.0) {
        double ratio = deltafu / deltaf;
        switch (scheme) {
        case OR_P2_PDM: {
            double x = one6th * (1.0 - 2.0 * cfl);
            limiter = (0.5 + x) + (0.5 - x) * ratio;
            limiter = dmin(dmin(2.0 * ratio / (cfl + 1.e-10), limiter),
                           2.0 / (1.0 - cfl));
            break;
        }
        case OR_SUPERBEE:
            limiter = dmax(dmin(2.0 * ratio, 1.0), dmin(ratio, 2.0));
            break;
        case OR_MUSCL:
            limiter = dmin(dmin(2.0 * ratio, 0.5 * (1.0 + ratio)), 2.0);
            break;
        case OR_P2:
            return reconstruction_p2(cfl, fu, deltafu, deltaf);
        case OR_UPSTREAM:
            return fu;
        default:
            break;
        }
        return fu + 0.5 * limiter * (1.0 - cfl) * deltaf;
    }
    if (scheme == OR_P2) return reconstruction_p2(cfl, fu, deltafu, deltaf);
    return fu;
}

/* ---- a8: masks ------------------------------------------------------- */
/* domain.F90:294-346 (serial: the halo updates of the masks are no-ops) */
int or_masks_from_az(const or_dom *d, const int32_t *az, int32_t *au,
                     int32_t *av, int32_t *ax)
{
    DOMVARS(d);
    memset(au, 0, sizeof(int32_t) * NXY);
    memset(av, 0, sizeof(int32_t) * NXY);
    memset(ax, 0, sizeof(int32_t) * NXY);
    for (int j = jl - HALO; j <= jh + HALO; j++)
        for (int i = il - HALO; i <= ih + HALO - 1; i++) {
            int a = az[IJ(i, j)], b = az[IJ(i + 1, j)];
            if (a == 1 && b == 1) au[IJ(i, j)] = 1;
            if ((a == 1 && b == 2) || (a == 2 && b == 1)) au[IJ(i, j)] = 2;
            if (a == 2 && b == 2) au[IJ(i, j)] = 3;
        }
    for (int j = jl - HALO; j <= jh + HALO - 1; j++)
        for (int i = il - HALO; i <= ih + HALO; i++) {
            int a = az[IJ(i, j)], b = az[IJ(i, j + 1)];
            if (a == 1 && b == 1) av[IJ(i, j)] = 1;
            if ((a == 1 && b == 2) || (a == 2 && b == 1)) av[IJ(i, j)] = 2;
            if (a == 2 && b == 2) av[IJ(i, j)] = 3;
        }
    for (int j = jl - HALO; j <= jh + HALO - 1; j++)
        for (int i = il - HALO; i <= ih + HALO - 1; i++)
            if (az[IJ(i, j)] >= 1 && az[IJ(i, j + 1)] >= 1 &&
                az[IJ(i + 1, j)] >= 1 && az[IJ(i + 1, j + 1)] >= 1)
                ax[IJ(i, j)] = 1;
    return OR_OK;
}

/* advection.F90:263-318.  The pointer sub-arrays of the reference are
 * stored here as full (NX,NY) arrays: a shifted view m(i,j) = src(i+1,j)
 * becomes a shifted copy whose never-addressed last row/column is 0. */
int or_init_advection(const or_dom *d, const int32_t *az, const int32_t *au,
                      const int32_t *av, const int32_t *ax, or_advgrids *g)
{
    DOMVARS(d);
    uint8_t *s = (uint8_t *)calloc((size_t)NXY * 9, 1);
    if (!s) return OR_ENOMEM;
    uint8_t *updH = s, *uflux = s + NXY, *vflux = s + 2 * NXY,
            *xflux = s + 3 * NXY, *uupdU = s + 4 * NXY, *vupdV = s + 5 * NXY,
            *ufluxU = s + 6 * NXY, *vfluxV = s + 7 * NXY;
    for (long n = 0; n < NXY; n++) {
        updH[n] = az[n] == 1;
        uflux[n] = au[n] == 1 || au[n] == 2;
        vflux[n] = av[n] == 1 || av[n] == 2;
        xflux[n] = ax[n] == 1;
        uupdU[n] = au[n] == 1;
        vupdV[n] = av[n] == 1;
    }
    for (int j = jl - HALO; j <= jh + HALO; j++) {
        uflux[IJ(ih + HALO, j)] = 0;
        xflux[IJ(ih + HALO, j)] = 0;
    }
    for (int i = il - HALO; i <= ih + HALO; i++) {
        vflux[IJ(i, jh + HALO)] = 0;
        xflux[IJ(i, jh + HALO)] = 0;
    }
    /* adv_gridU%mask_uflux = mask_updateH(1+_IRANGE_HALO_,...) */
    for (int j = jl - HALO; j <= jh + HALO; j++)
        for (int i = il - HALO; i <= ih + HALO - 1; i++)
            ufluxU[IJ(i, j)] = updH[IJ(i + 1, j)];
    /* adv_gridV%mask_vflux => mask_updateH(_IRANGE_HALO_,1+_JRANGE_HALO_) */
    for (int j = jl - HALO; j <= jh + HALO - 1; j++)
        for (int i = il - HALO; i <= ih + HALO; i++)
            vfluxV[IJ(i, j)] = updH[IJ(i, j + 1)];

    g->store = s;
    g->H.mask_uflux = uflux;   g->H.mask_vflux = vflux;
    g->H.mask_uupdate = updH;  g->H.mask_vupdate = updH;
    g->H.mask_finalise = updH; g->H.az = (int32_t *)az;
    g->U.mask_uflux = ufluxU;  g->U.mask_vflux = xflux;
    g->U.mask_uupdate = uupdU; g->U.mask_vupdate = uflux;
    g->U.mask_finalise = uflux; g->U.az = (int32_t *)au;
    g->V.mask_uflux = xflux;   g->V.mask_vflux = vfluxV;
    g->V.mask_uupdate = vflux; g->V.mask_vupdate = vupdV;
    g->V.mask_finalise = vflux; g->V.az = (int32_t *)av;
    return OR_OK;
}

void or_free_advection(or_advgrids *g)
{
    free(g->store);
    memset(g, 0, sizeof(*g));
}

/* ---- a2: adv_split_u.F90:205-286 -------------------------------------- */
void or_adv_split_u(const or_dom *d, double dt, const double *f, double *fi,
                    double *Di, double *adv, const double *U,
                    const double *DU, double splitfac, int scheme, double AH,
                    const uint8_t *mask_flux, const uint8_t *mask_update)
{
    DOMVARS(d);
    const double dx = d->dx, dy = d->dy, ard1 = d->ard1;
    const int isub = (scheme == OR_UPSTREAM) ? 0 : 1;
    const int use_AH = AH > 0.0;
    const double dti = splitfac * dt;
    double *uflux = scr(SCR_FLUX, (size_t)NXY);

#pragma omp parallel OMP_NT
    {
#pragma omp for schedule(static)
        for (int j = jl - HALO; j <= jh + HALO; j++) {
            for (int i = il - HALO + isub; i <= ih + HALO - 1 - isub; i++) {
                long c = IJ(i, j);
                if (mask_flux[c]) {
                    int use_limiter = 0;
                    double fu, fuu = 0, fd = 0, cfl = 0;
                    if (U[c] > 0.0) {
                        fu = f[c];
                        if (scheme != OR_UPSTREAM)
                            use_limiter = mask_flux[IJ(i - 1, j)];
                        if (use_limiter) {
                            cfl = U[c] / DU[c] * dti / dx;
                            fuu = f[IJ(i - 1, j)];
                            fd = f[IJ(i + 1, j)];
                        }
                    } else {
                        fu = f[IJ(i + 1, j)];
                        if (scheme != OR_UPSTREAM)
                            use_limiter = mask_flux[IJ(i + 1, j)];
                        if (use_limiter) {
                            cfl = -U[c] / DU[c] * dti / dx;
                            fuu = f[IJ(i + 2, j)];
                            fd = f[c];
                        }
                    }
                    if (use_limiter)
                        fu = or_adv_interfacial_reconstruction(scheme, cfl,
                                                               fuu, fu, fd);
                    uflux[c] = U[c] * fu;
                    if (use_AH)
                        uflux[c] = uflux[c] -
                                   AH * DU[c] * (f[IJ(i + 1, j)] - f[c]) / dx;
                } else {
                    uflux[c] = 0.0;
                }
            }
        }
#pragma omp for schedule(static)
        for (int j = jl - HALO; j <= jh + HALO; j++) {
            for (int i = il - HALO + 1 + isub; i <= ih + HALO - 1 - isub; i++) {
                long c = IJ(i, j), w = IJ(i - 1, j);
                if (mask_update[c]) {
                    double Dio = Di[c];
                    Di[c] = Dio - dti * (U[c] * dy - U[w] * dy) * ard1;
                    double advn = splitfac * (uflux[c] * dy - uflux[w] * dy) * ard1;
                    fi[c] = (Dio * fi[c] - dt * advn) / Di[c];
                    adv[c] = adv[c] + advn;
                }
            }
        }
    }
}

/* ---- a3: adv_split_v.F90:63-136 --------------------------------------- */
void or_adv_split_v(const or_dom *d, double dt, const double *f, double *fi,
                    double *Di, double *adv, const double *V,
                    const double *DV, double splitfac, int scheme, double AH,
                    const uint8_t *mask_flux, const uint8_t *mask_update)
{
    DOMVARS(d);
    const double dx = d->dx, dy = d->dy, ard1 = d->ard1;
    const int jsub = (scheme == OR_UPSTREAM) ? 0 : 1;
    const int use_AH = AH > 0.0;
    const double dti = splitfac * dt;
    double *vflux = scr(SCR_FLUX, (size_t)NXY);

#pragma omp parallel OMP_NT
    {
#pragma omp for schedule(static)
        for (int j = jl - HALO + jsub; j <= jh + HALO - 1 - jsub; j++) {
            for (int i = il - HALO; i <= ih + HALO; i++) {
                long c = IJ(i, j);
                if (mask_flux[c]) {
                    int use_limiter = 0;
                    double fu, fuu = 0, fd = 0, cfl = 0;
                    if (V[c] > 0.0) {
                        fu = f[c];
                        if (scheme != OR_UPSTREAM)
                            use_limiter = mask_flux[IJ(i, j - 1)];
                        if (use_limiter) {
                            cfl = V[c] / DV[c] * dti / dy;
                            fuu = f[IJ(i, j - 1)];
                            fd = f[IJ(i, j + 1)];
                        }
                    } else {
                        fu = f[IJ(i, j + 1)];
                        if (scheme != OR_UPSTREAM)
                            use_limiter = mask_flux[IJ(i, j + 1)];
                        if (use_limiter) {
                            cfl = -V[c] / DV[c] * dti / dy;
                            fuu = f[IJ(i, j + 2)];
                            fd = f[c];
                        }
                    }
                    if (use_limiter)
                        fu = or_adv_interfacial_reconstruction(scheme, cfl,
                                                               fuu, fu, fd);
                    vflux[c] = V[c] * fu;
                    if (use_AH)
                        vflux[c] = vflux[c] -
                                   AH * DV[c] * (f[IJ(i, j + 1)] - f[c]) / dy;
                } else {
                    vflux[c] = 0.0;
                }
            }
        }
#pragma omp for schedule(static)
        for (int j = jl - HALO + 1 + jsub; j <= jh + HALO - 1 - jsub; j++) {
            for (int i = il - HALO; i <= ih + HALO; i++) {
                long c = IJ(i, j), s = IJ(i, j - 1);
                if (mask_update[c]) {
                    double Dio = Di[c];
                    Di[c] = Dio - dti * (V[c] * dx - V[s] * dx) * ard1;
                    double advn = splitfac * (vflux[c] * dx - vflux[s] * dx) * ard1;
                    fi[c] = (Dio * fi[c] - dt * advn) / Di[c];
                    adv[c] = adv[c] + advn;
                }
            }
        }
    }
}

/* ---- a4: adv_split_w.F90:85-353 (the non-_POINTER_REMAP_ path) -------- */
long or_adv_split_w(const or_dom *d, double dt, const double *f, double *fi,
                    double *hi, double *adv, const double *ww,
                    double splitfac, int scheme, int tag, const int32_t *az,
                    int itersmax)
{
    DOMVARS(d);
    if (scheme == OR_NOADV) return 0;
    const int kshift = (tag == OR_W_TAG) ? 1 : 0;
    const double dti = splitfac * dt;
    const int iterate = itersmax > 1;
    long nwarn = 0;

#pragma omp parallel OMP_NT reduction(+ : nwarn)
    {
        double *wflux = (double *)malloc(sizeof(double) * (kmax + 1) * 8);
        double *fo = wflux + (kmax + 1), *fiaux = fo + (kmax + 1),
               *hiaux = fiaux + (kmax + 1), *advaux = hiaux + (kmax + 1),
               *cfl0 = advaux + (kmax + 1), *cflsb = cfl0 + (kmax + 1);
        double *cfls = iterate ? cflsb : cfl0;
        wflux[0] = 0.0;
        wflux[kmax] = 0.0;
#pragma omp for schedule(static)
        for (int j = jl - HALO; j <= jh + HALO - 1; j++) {
            for (int i = il - HALO; i <= ih + HALO - 1; i++) {
                if (az[IJ(i, j)] != 1) continue;
                const long c0 = IJ(i, j);
#define COL(a, k) (a)[c0 + NXY * (long)(k)]
                if (scheme != OR_UPSTREAM || iterate)
                    for (int k = 1 - kshift; k <= kmax - 1; k++)
                        cfl0[k] = fabs(COL(ww, k)) * dti /
                                  (0.5 * (COL(hi, k) + COL(hi, k + 1)));
                int iters = 1;
                double itersm1 = 1.0, dtik = dti, splitfack = splitfac;
                for (int k = 0; k <= kmax; k++) fo[k] = COL(f, k);
                const double *faux = fo;
                int it = 1;
                while (it <= iters) {
                    if (it == 1) {
                        for (int k = 0; k <= kmax; k++) {
                            fiaux[k] = COL(fi, k);
                            hiaux[k] = COL(hi, k);
                            advaux[k] = COL(adv, k);
                        }
                        faux = fo;
                    }
                    if (iterate) {
                        if (it == 1) {
                            for (int k = 0; k <= kmax; k++) cfls[k] = cfl0[k];
                        } else if (scheme != OR_UPSTREAM || iters < itersmax) {
                            for (int k = 1 - kshift; k <= kmax - 1; k++)
                                cfls[k] = fabs(COL(ww, k)) * dti /
                                          (0.5 * (hiaux[k] + hiaux[k + 1]));
                        }
                        if (iters < itersmax) {
                            int mx = -2147483647 - 1;
                            for (int k = 1 - kshift; k <= kmax - 1; k++) {
                                double cc = ceil(cfls[k]);
                                int ci = cc > 2147483647.0 ? 2147483647 : (int)cc;
                                if (ci > mx) mx = ci;
                            }
                            int iters_new = mx > 1 ? mx : 1;
                            if (iters_new > iters) {
                                if (iters_new > itersmax) {
                                    nwarn++;
                                    iters = itersmax;
                                } else {
                                    iters = iters_new;
                                }
                                itersm1 = 1.0 / iters;
                                dtik = dti * itersm1;
                                splitfack = splitfac * itersm1;
                                if (it > 1) {
                                    it = 1;
                                    continue;
                                }
                            }
                        }
                    }
                    for (int k = 1 - kshift; k <= kmax - 1; k++) {
                        int use_limiter = 0;
                        double fu, fuu = 0, fd = 0;
                        if (COL(ww, k) > 0.0) {
                            fu = faux[k];
                            if (scheme != OR_UPSTREAM) use_limiter = k > 1 - kshift;
                            if (use_limiter) {
                                fuu = faux[k - 1];
                                fd = faux[k + 1];
                            }
                        } else {
                            fu = faux[k + 1];
                            if (scheme != OR_UPSTREAM) use_limiter = k < kmax - 1;
                            if (use_limiter) {
                                fuu = faux[k + 2];
                                fd = faux[k];
                            }
                        }
                        if (use_limiter)
                            fu = or_adv_interfacial_reconstruction(
                                scheme, cfls[k] * itersm1, fuu, fu, fd);
                        wflux[k] = COL(ww, k) * fu;
                    }
                    for (int k = 1; k <= kmax - kshift; k++) {
                        double hio = hiaux[k];
                        hiaux[k] = hio - dtik * (COL(ww, k) - COL(ww, k - 1));
                        double advn = splitfack * (wflux[k] - wflux[k - 1]);
                        fiaux[k] = (hio * fiaux[k] - dt * advn) / hiaux[k];
                        advaux[k] = advaux[k] + advn;
                    }
                    faux = fiaux;
                    it++;
                }
                for (int k = 1; k <= kmax - kshift; k++) {
                    COL(fi, k) = fiaux[k];
                    COL(hi, k) = hiaux[k];
                    COL(adv, k) = advaux[k];
                }
#undef COL
            }
        }
        free(wflux);
    }
    return nwarn;
}

/* ---- a7: adv_upstream_2dh.F90:70-160 ---------------------------------- */
void or_adv_upstream_2dh(const or_dom *d, double dt, const double *f,
                         double *fi, double *Di, double *adv,
                         const double *U, const double *V, const double *Dn,
                         const double *DU, const double *DV, double AH,
                         const int32_t *az)
{
    DOMVARS(d);
    const double dx = d->dx, dy = d->dy, ard1 = d->ard1;
    double *uflux = scr(SCR_FLUX, (size_t)NXY);
    double *vflux = scr(SCR_VFLUX, (size_t)NXY);
    double *cmin = scr(SCR_CMIN, (size_t)NXY);
    double *cmax = scr(SCR_CMAX, (size_t)NXY);
    (void)DU; (void)DV;

#pragma omp parallel OMP_NT
    {
#pragma omp for schedule(static) nowait
        for (int j = jl; j <= jh; j++)
            for (int i = il - 1; i <= ih; i++) {
                long c = IJ(i, j), e = IJ(i + 1, j);
                if (U[c] > 0.0) uflux[c] = U[c] * f[c];
                else            uflux[c] = U[c] * f[e];
                if (AH > 0.0 && az[c] > 0 && az[e] > 0)
                    uflux[c] = uflux[c] -
                               AH * (f[e] - f[c]) / dx * 0.5 * (Dn[e] + Dn[c]);
            }
#pragma omp for schedule(static) nowait
        for (int j = jl - 1; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                long c = IJ(i, j), n = IJ(i, j + 1);
                if (V[c] > 0.0) vflux[c] = V[c] * f[c];
                else            vflux[c] = V[c] * f[n];
                if (AH > 0.0 && az[c] > 0 && az[n] > 0)
                    vflux[c] = vflux[c] -
                               AH * (f[n] - f[c]) / dy * 0.5 * (Dn[n] + Dn[c]);
            }
#pragma omp for schedule(static)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                long c = IJ(i, j);
                if (az[c] == 1) {
                    double mn = 100000 * 1.0, mx = -100000 * 1.0;
                    for (int ii = -1; ii <= 1; ii++) {
                        long q = IJ(i + ii, j);
                        if (az[q] == 1) {
                            if (f[q] > mx) mx = f[q];
                            if (f[q] < mn) mn = f[q];
                        }
                    }
                    for (int jj = -1; jj <= 1; jj += 2) {
                        long q = IJ(i, j + jj);
                        if (az[q] == 1) {
                            if (f[q] > mx) mx = f[q];
                            if (f[q] < mn) mn = f[q];
                        }
                    }
                    cmin[c] = mn;
                    cmax[c] = mx;
                }
            }
#pragma omp for schedule(static)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                long c = IJ(i, j), w = IJ(i - 1, j), s = IJ(i, j - 1);
                if (az[c] == 1) {
                    double Dio = Di[c];
                    Di[c] = Dio - dt * (U[c] * dy - U[w] * dy +
                                        V[c] * dx - V[s] * dx) * ard1;
                    double advn = (uflux[c] * dy - uflux[w] * dy +
                                   vflux[c] * dx - vflux[s] * dx) * ard1;
                    fi[c] = (Dio * fi[c] - dt * advn) / Di[c];
                    adv[c] = adv[c] + advn;
                    if (fi[c] > cmax[c]) fi[c] = cmax[c];
                    if (fi[c] < cmin[c]) fi[c] = cmin[c];
                }
            }
    }
}

/* ---- a6: do_advection, advection.F90:362-702 --------------------------- */
static int split_scheme(int s)
{
    return s == OR_UPSTREAM || s == OR_P2 || s == OR_SUPERBEE ||
           s == OR_MUSCL || s == OR_P2_PDM;
}

static const or_advgrid *grid_for_tag(const or_advgrids *g, int tag)
{
    switch (tag) {
    case OR_H_TAG: case OR_D_TAG: case OR_W_TAG: return &g->H;
    case OR_U_TAG: return &g->U;
    case OR_V_TAG: return &g->V;
    default: return NULL;
    }
}

int or_do_advection(const or_dom *d, const or_advgrids *g, double dt,
                    double *f, const double *U, const double *V,
                    const double *DU, const double *DV, const double *Do,
                    const double *Dn, int split, int scheme, double AH,
                    int tag, double *Dires, double *advres)
{
    DOMVARS(d);
    if (tag == OR_W_TAG) return OR_EINVAL_TAG;  /* 2-D driver: H,D,U,V only */
    const or_advgrid *ag = grid_for_tag(g, tag);
    if (!ag) return OR_EINVAL_TAG;
    /* automatic fi/Di/adv of the reference; slots distinct from the 3-D
     * driver's, which may be passing its own slices in as Dires/advres */
    static double *own = NULL;
    static long own_n = 0;
    if (own_n < 3 * NXY) {
        free(own);
        own = (double *)malloc(sizeof(double) * 3 * NXY);
        own_n = own ? 3 * NXY : 0;
        if (!own) return OR_ENOMEM;
    }
    double *fi = own, *pDi = Dires ? Dires : own + NXY,
           *padv = advres ? advres : own + 2 * NXY;
    memcpy(pDi, Do, sizeof(double) * NXY);
    memset(padv, 0, sizeof(double) * NXY);
    if (scheme == OR_NOADV) return OR_OK;

    switch (split) {
    case OR_NOSPLIT:
        if (split_scheme(scheme)) {
            memcpy(fi, f, sizeof(double) * NXY);
            or_adv_split_u(d, dt, f, fi, pDi, padv, U, DU, 1.0, scheme, AH,
                           ag->mask_uflux, ag->mask_uupdate);
            or_adv_split_v(d, dt, f, fi, pDi, padv, V, DV, 1.0, scheme, AH,
                           ag->mask_vflux, ag->mask_vupdate);
            for (int j = jl - HALO; j <= jh + HALO; j++)
                for (int i = il - HALO; i <= ih + HALO; i++) {
                    long c = IJ(i, j);
                    if (ag->mask_finalise[c])
                        f[c] = (Do[c] * f[c] - dt * padv[c]) / pDi[c];
                }
        } else if (scheme == OR_UPSTREAM_2DH) {
            or_adv_upstream_2dh(d, dt, f, f, pDi, padv, U, V, Dn, DU, DV, AH,
                                ag->az);
        } else if (scheme == OR_J7 || scheme == OR_FCT || scheme == OR_P2_2DH) {
            return OR_ESCHEME_SPLIT;  /* not part of this build */
        } else {
            return OR_EINVAL_SCHEME;
        }
        break;
    case OR_FULLSPLIT:
        if (split_scheme(scheme)) {
            or_adv_split_u(d, dt, f, f, pDi, padv, U, DU, 1.0, scheme, AH,
                           ag->mask_uflux, ag->mask_uupdate);
            or_adv_split_v(d, dt, f, f, pDi, padv, V, DV, 1.0, scheme, AH,
                           ag->mask_vflux, ag->mask_vupdate);
        } else if (scheme == OR_UPSTREAM_2DH || scheme == OR_J7 ||
                   scheme == OR_FCT || scheme == OR_P2_2DH) {
            return OR_ESCHEME_SPLIT;
        } else {
            return OR_EINVAL_SCHEME;
        }
        break;
    case OR_HALFSPLIT:
        if (split_scheme(scheme)) {
            or_adv_split_u(d, dt, f, f, pDi, padv, U, DU, 0.5, scheme, AH,
                           ag->mask_uflux, ag->mask_uupdate);
            or_adv_split_v(d, dt, f, f, pDi, padv, V, DV, 1.0, scheme, AH,
                           ag->mask_vflux, ag->mask_vupdate);
            or_adv_split_u(d, dt, f, f, pDi, padv, U, DU, 0.5, scheme, AH,
                           ag->mask_uflux, ag->mask_uupdate);
        } else if (scheme == OR_UPSTREAM_2DH || scheme == OR_J7 ||
                   scheme == OR_FCT || scheme == OR_P2_2DH) {
            return OR_ESCHEME_SPLIT;
        } else {
            return OR_EINVAL_SCHEME;
        }
        break;
    default:
        return OR_EINVAL_SPLIT;
    }
    return OR_OK;
}

/* ---- a5: do_advection_3d, advection_3d.F90:120-598 --------------------- */
int or_do_advection_3d(const or_dom *d, const or_advgrids *g, double dt,
                       double *f, const double *uu, const double *vv,
                       const double *ww, const double *hu, const double *hv,
                       const double *ho, const double *hn, int split,
                       int hscheme, int vscheme, double AH, int tag,
                       int adv_ver_iterations, double *hires, double *advres,
                       long *nwarn)
{
    DOMVARS(d);
    int tag2d;
    const or_advgrid *ag;
    long nw = 0;
    switch (tag) {
    case OR_H_TAG: case OR_D_TAG: tag2d = OR_H_TAG; ag = &g->H; break;
    case OR_U_TAG: tag2d = OR_U_TAG; ag = &g->U; break;
    case OR_V_TAG: tag2d = OR_V_TAG; ag = &g->V; break;
    case OR_W_TAG: tag2d = OR_H_TAG; ag = &g->H; break;
    default: return OR_EINVAL_TAG;
    }
    const long N3 = NXY * (kmax + 1);
    double *fi = scr(SCR_FI, (size_t)N3);
    double *phi = hires ? hires : scr(SCR_HI, (size_t)N3);
    double *padv = advres ? advres : scr(SCR_ADV, (size_t)N3);
    if (!fi || !phi || !padv) return OR_ENOMEM;
    memcpy(phi, ho, sizeof(double) * N3);
    memset(padv, 0, sizeof(double) * N3);
    int rc = OR_OK;

    if (hscheme != OR_NOADV || vscheme != OR_NOADV) {
        if (hscheme == OR_NOADV) {
            if (kmax > 1)
                nw += or_adv_split_w(d, dt, f, f, phi, padv, ww, 1.0, vscheme,
                                     tag, ag->az, adv_ver_iterations);
        } else if (vscheme == OR_NOADV) {
            for (int k = 1; k <= kmax && rc == OR_OK; k++) {
                long o = NXY * k;
                rc = or_do_advection(d, g, dt, f + o, uu + o, vv + o, hu + o,
                                     hv + o, ho + o, hn + o, split, hscheme,
                                     AH, tag2d, phi + o, padv + o);
            }
        } else {
            switch (split) {
            case OR_NOSPLIT:
                memcpy(fi, f, sizeof(double) * N3);
                if (split_scheme(hscheme)) {
                    for (int k = 1; k <= kmax; k++) {
                        long o = NXY * k;
                        or_adv_split_u(d, dt, f + o, fi + o, phi + o, padv + o,
                                       uu + o, hu + o, 1.0, hscheme, AH,
                                       ag->mask_uflux, ag->mask_uupdate);
                        or_adv_split_v(d, dt, f + o, fi + o, phi + o, padv + o,
                                       vv + o, hv + o, 1.0, hscheme, AH,
                                       ag->mask_vflux, ag->mask_vupdate);
                    }
                } else if (hscheme == OR_UPSTREAM_2DH) {
                    for (int k = 1; k <= kmax; k++) {
                        long o = NXY * k;
                        or_adv_upstream_2dh(d, dt, f + o, fi + o, phi + o,
                                            padv + o, uu + o, vv + o, hn + o,
                                            hu + o, hv + o, AH, ag->az);
                    }
                } else if (hscheme == OR_J7 || hscheme == OR_FCT ||
                           hscheme == OR_P2_2DH) {
                    return OR_ESCHEME_SPLIT;  /* not part of this build */
                } else {
                    return OR_EINVAL_SCHEME;
                }
                if (kmax > 1)
                    nw += or_adv_split_w(d, dt, f, fi, phi, padv, ww, 1.0,
                                         vscheme, tag, ag->az, 1);
                for (int j = jl - HALO; j <= jh + HALO; j++)
                    for (int i = il - HALO; i <= ih + HALO; i++) {
                        if (!ag->mask_finalise[IJ(i, j)]) continue;
                        for (int k = 1; k <= kmax; k++) {
                            long c = IJK(i, j, k);
                            f[c] = (ho[c] * f[c] - dt * padv[c]) / phi[c];
                        }
                    }
                break;
            case OR_FULLSPLIT:
                for (int k = 1; k <= kmax && rc == OR_OK; k++) {
                    long o = NXY * k;
                    rc = or_do_advection(d, g, dt, f + o, uu + o, vv + o,
                                         hu + o, hv + o, ho + o, hn + o,
                                         OR_FULLSPLIT, hscheme, AH, tag2d,
                                         phi + o, padv + o);
                }
                if (rc != OR_OK) break;
                if (kmax > 1)
                    nw += or_adv_split_w(d, dt, f, f, phi, padv, ww, 1.0,
                                         vscheme, tag, ag->az,
                                         adv_ver_iterations);
                break;
            case OR_HALFSPLIT:
                if (!split_scheme(hscheme)) {
                    if (hscheme == OR_UPSTREAM_2DH || hscheme == OR_J7 ||
                        hscheme == OR_FCT || hscheme == OR_P2_2DH)
                        return OR_ESCHEME_SPLIT;
                    return OR_EINVAL_SCHEME;
                }
                for (int k = 1; k <= kmax; k++) {
                    long o = NXY * k;
                    or_adv_split_u(d, dt, f + o, f + o, phi + o, padv + o,
                                   uu + o, hu + o, 0.5, hscheme, AH,
                                   ag->mask_uflux, ag->mask_uupdate);
                }
                for (int k = 1; k <= kmax; k++) {
                    long o = NXY * k;
                    or_adv_split_v(d, dt, f + o, f + o, phi + o, padv + o,
                                   vv + o, hv + o, 0.5, hscheme, AH,
                                   ag->mask_vflux, ag->mask_vupdate);
                }
                if (kmax > 1)
                    nw += or_adv_split_w(d, dt, f, f, phi, padv, ww, 1.0,
                                         vscheme, tag, ag->az,
                                         adv_ver_iterations);
                for (int k = 1; k <= kmax; k++) {
                    long o = NXY * k;
                    or_adv_split_v(d, dt, f + o, f + o, phi + o, padv + o,
                                   vv + o, hv + o, 0.5, hscheme, AH,
                                   ag->mask_vflux, ag->mask_vupdate);
                }
                for (int k = 1; k <= kmax; k++) {
                    long o = NXY * k;
                    or_adv_split_u(d, dt, f + o, f + o, phi + o, padv + o,
                                   uu + o, hu + o, 0.5, hscheme, AH,
                                   ag->mask_uflux, ag->mask_uupdate);
                }
                break;
            case OR_HVSPLIT:
                for (int k = 1; k <= kmax && rc == OR_OK; k++) {
                    long o = NXY * k;
                    rc = or_do_advection(d, g, dt, f + o, uu + o, vv + o,
                                         hu + o, hv + o, ho + o, hn + o,
                                         OR_NOSPLIT, hscheme, AH, tag2d,
                                         phi + o, padv + o);
                }
                if (rc != OR_OK) break;
                if (kmax > 1)
                    nw += or_adv_split_w(d, dt, f, f, phi, padv, ww, 1.0,
                                         vscheme, tag, ag->az,
                                         adv_ver_iterations);
                break;
            default:
                return OR_EINVAL_SPLIT;
            }
        }
    }
    if (nwarn) *nwarn = nw;
    return rc;
}

/* ---- a12: eqstate.F90:127-149 (case 1) and :252-261 --------------------- */
void or_eos_linear(const or_dom *d, const double *T, const double *S,
                   double *rho, double *buoy, double T0, double S0,
                   double dtr0, double dsr0)
{
    DOMVARS(d);
    const double g = (double)9.81f, rho_0 = (double)1025.f;
    const double x = -g / rho_0;
#pragma omp parallel for OMP_NT schedule(static) collapse(2)
    for (int k = 1; k <= kmax; k++)
        for (int j = jl - HALO; j <= jh + HALO; j++)
            for (int i = il - HALO; i <= ih + HALO; i++) {
                long c = IJK(i, j, k);
                rho[c] = rho_0 + dtr0 * (T[c] - T0) + dsr0 * (S[c] - S0);
                buoy[c] = x * (rho[c] - rho_0);
            }
}

/* ---- full equation of state, eqstate.F90:94-562 --------------------------
 * Literals written without a kind in the reference are default REAL (FP32)
 * and enter the FP64 arithmetic rounded to FP32; the Jackett et al. (2006)
 * coefficients are d0 literals (FP64).  Every expression keeps the
 * reference's left-to-right order. */

/* rho_from_theta_unesco80, eqstate.F90:280-329 (UNESCO 1980 one-atmosphere
 * density; NONNEGSALT clamps negative salinity, :309-317) */
double or_rho_unesco80(double T, double S, int nonnegsalt)
{
    const double t1 = T, t2 = t1 * t1, t3 = t1 * t2, t4 = t2 * t2,
                 t5 = t1 * t4;
    double s1 = S;
    if (nonnegsalt && s1 < 0.0) s1 = 0.0;
    const double s2 = s1 * s1, s3 = s1 * s2;
    double x = 999.842594f + 6.793952e-02f * t1 - 9.09529e-03f * t2 +
               1.001685e-04f * t3;
    x = x - 1.120083e-06f * t4 + 6.536332e-09f * t5;
    x = x + s1 * (0.824493f - 4.0899e-03f * t1 + 7.6438e-05f * t2 -
                  8.2467e-07f * t3);
    x = x + s1 * 5.3875e-09f * t4;
    x = x + sqrt(s3) * (-5.72466e-03f + 1.0227e-04f * t1 - 1.6546e-06f * t2);
    x = x + 4.8314e-04f * s2;
    return x;
}

/* numerator and denominator of the Jackett et al. (2006) rational function
 * at p = 0, eqstate.F90:381-400 (same terms at :478-498) */
static void jackett_p0(double s, double th, double th2, double sqrts,
                       double *anum, double *aden)
{
    *anum = 9.9984085444849347e+02 +
            th * (7.3471625860981584e+00 +
                  th * (-5.3211231792841769e-02 + th * 3.6492439109814549e-04)) +
            s * (2.5880571023991390e+00 - th * 6.7168282786692355e-03 +
                 s * 1.9203202055760151e-03);
    *aden = 1.0000000000000000e+00 +
            th * (7.2815210113327091e-03 +
                  th * (-4.4787265461983921e-05 +
                        th * (3.3851002965802430e-07 +
                              th * 1.3651202389758572e-10))) +
            s * (1.7632126669040377e-03 -
                 th * (8.8066583251206474e-06 + th2 * 1.8832689434804897e-10) +
                 sqrts * (5.7463776745432097e-06 + th2 * 1.4716275472242334e-09));
}

/* the pressure terms, eqstate.F90:407-417 (same at :526-536) */
static void jackett_p(double s, double th2, double p, double pth, double *anum,
                      double *aden)
{
    *anum = *anum + p * (1.1798263740430364e-02 + th2 * 9.8920219266399117e-08 +
                         s * 4.6996642771754730e-06 -
                         p * (2.5862187075154352e-08 + th2 * 3.2921414007960662e-12));
    *aden = *aden + p * (6.7103246285651894e-06 -
                         pth * (th2 * 2.4461698007024582e-17 +
                                p * 9.1534417604289062e-18));
}

/* rho_from_theta, eqstate.F90:338-423 */
void or_rho_from_theta(double s, double th, double p, double *dens0,
                       double *densp)
{
    const double th2 = th * th, sqrts = sqrt(s);
    double anum, aden;
    jackett_p0(s, th, th2, sqrts, &anum, &aden);
    *dens0 = anum / aden;
    *densp = *dens0;
    if (p != 0.0) {
        jackett_p(s, th2, p, p * th, &anum, &aden);
        *densp = anum / aden;
    }
}

/* eosall_from_theta, eqstate.F90:432-562: d(rho)/ds and d(rho)/d(theta) */
void or_eosall_from_theta(double s, double th, double p, double *rho_s,
                          double *rho_th)
{
    const double th2 = th * th, sqrts = sqrt(s);
    double anum, aden;
    jackett_p0(s, th, th2, sqrts, &anum, &aden);
    double anum_s = 2.5880571023991390e+00 - th * 6.7168282786692355e-03 +
                    s * 3.8406404111520300e-03;
    const double aden_s =
        1.7632126669040377e-03 +
        th * (-8.8066583251206470e-06 - th2 * 1.8832689434804897e-10) +
        sqrts * (8.6195665118148150e-06 + th2 * 2.2074413208363504e-09);
    double anum_th = 7.3471625860981580e+00 +
                     th * (-1.0642246358568354e-01 + th * 1.0947731732944364e-03) -
                     s * 6.7168282786692355e-03;
    double aden_th =
        7.2815210113327090e-03 +
        th * (-8.9574530923967840e-05 +
              th * (1.0155300889740728e-06 + th * 5.4604809559034290e-10)) +
        s * (-8.8066583251206470e-06 - th2 * 5.6498068304414700e-10 +
             th * sqrts * 2.9432550944484670e-09);
    if (p != 0.0) {
        const double pth = p * th;
        jackett_p(s, th2, p, pth, &anum, &aden);
        anum_s = anum_s + p * 4.6996642771754730e-06;
        anum_th = anum_th + pth * (1.9784043853279823e-07 - p * 6.5842828015921320e-12);
        aden_th = aden_th - p * p * (th2 * 7.3385094021073750e-17 +
                                     p * 9.1534417604289060e-18);
    }
    const double rec_aden = 1.0 / aden;
    const double rho = anum * rec_aden;
    *rho_s = (anum_s - aden_s * rho) * rec_aden;
    *rho_th = (anum_th - aden_th * rho) * rec_aden;
}

/* do_eqstate, eqstate.F90:94-271, with kmin = 1 (sigma and general
 * coordinates, sigma_coordinates.F90:78, general_coordinates.F90:104).
 * method 1: linear rho on every cell, alpha = dtr0 and beta = dsr0 on the
 * whole array (:127-149); 2: UNESCO 1980 and 3: Jackett et al. (2006) on
 * water columns (az > 0), land keeps its rho (:151-219); for 2 and 3 the
 * expansion coefficients at the layer interfaces from the interface means of
 * T and S and the pressure accumulated from the top (:168-190, :222-245).
 * Then buoy on every cell from rho (:252-261).  alpha/beta may be NULL (the
 * _OLD_BVF_ build does not compute them).  Returns 0, or 1 for a bad method
 * (init_eqstate stops, :75-77). */
int or_do_eqstate(const or_dom *d, const int32_t *az, int method,
                  int nonnegsalt, const double *T, const double *S,
                  const double *hn, double *rho, double *buoy, double *alpha,
                  double *beta, double T0, double S0, double dtr0, double dsr0)
{
    DOMVARS(d);
    if (method < 1 || method > 3) return 1;
    const double g = (double)9.81f, rho_0 = (double)1025.f;
    if (method == 1) {
        for (long k = 1; k <= kmax; k++)
            for (long c = 0; c < NXY; c++) {
                const long q = c + NXY * k;
                rho[q] = rho_0 + dtr0 * (T[q] - T0) + dsr0 * (S[q] - S0);
            }
        if (alpha)
            for (long q = 0; q < NXY * (kmax + 1); q++) alpha[q] = dtr0;
        if (beta)
            for (long q = 0; q < NXY * (kmax + 1); q++) beta[q] = dsr0;
    } else {
#pragma omp parallel for OMP_NT schedule(static)
        for (long c = 0; c < NXY; c++) {
            if (az[c] <= 0) continue;
            for (int k = 1; k <= kmax; k++) {
                const long q = c + NXY * k;
                if (method == 2) {
                    rho[q] = or_rho_unesco80(T[q], S[q], nonnegsalt);
                } else {
                    double s1 = S[q], densp;
                    if (nonnegsalt && s1 < 0.0) s1 = 0.0;
                    or_rho_from_theta(s1, T[q], 0.0, &rho[q], &densp);
                }
            }
            if (alpha && beta) {
                alpha[c + NXY * kmax] = 0.0;
                beta[c + NXY * kmax] = 0.0;
                double p1 = 0.0;
                for (int k = kmax - 1; k >= 1; k--) {
                    const long q = c + NXY * k, qu = q + NXY;
                    const double th = 0.5 * (T[qu] + T[q]);
                    const double s1 = 0.5 * (S[qu] + S[q]);
                    p1 = p1 + hn[qu];
                    or_eosall_from_theta(s1, th, p1, &beta[q], &alpha[q]);
                }
            }
        }
    }
    const double x = -g / rho_0;
    for (long q = NXY; q < NXY * (kmax + 1); q++) buoy[q] = x * (rho[q] - rho_0);
    return 0;
}

/* ---- a9: ip_blumberg_mellor.F90:67-150 ---------------------------------- */
void or_ip_blumberg_mellor(const or_dom *d, const int32_t *az,
                           const int32_t *au, const int32_t *av,
                           const double *H, const double *hn,
                           const double *hun, const double *hvn,
                           const double *buoy, double *zz, double *idpdx,
                           double *idpdy)
{
    DOMVARS(d);
    const double dxm1 = 1.0 / d->dx, dym1 = 1.0 / d->dy;
    for (long c = 0; c < NXY; c++) {
        zz[c] = 0.0;
        idpdx[c] = 0.0;
        idpdy[c] = 0.0;
    }
#pragma omp parallel OMP_NT
    {
#pragma omp for schedule(static)
        for (int j = jl; j <= jh + 1; j++)
            for (int i = il; i <= ih + 1; i++)
                if (az[IJ(i, j)] >= 1) {
                    zz[IJK(i, j, 1)] = -H[IJ(i, j)] + hn[IJK(i, j, 1)];
                    for (int k = 2; k <= kmax; k++)
                        zz[IJK(i, j, k)] = zz[IJK(i, j, k - 1)] + hn[IJK(i, j, k)];
                }
#pragma omp for schedule(static)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++)
                if (au[IJ(i, j)] >= 1) {
#define B(ii, k) buoy[IJK(ii, j, k)]
                    double grdl = (B(i + 1, kmax) - B(i, kmax)) * dxm1;
                    double buoyl = 0.5 * (B(i + 1, kmax) + B(i, kmax));
                    double prgr = grdl * 0.5 * hun[IJK(i, j, kmax)];
                    idpdx[IJK(i, j, kmax)] = hun[IJK(i, j, kmax)] * prgr;
                    for (int k = kmax - 1; k >= 1; k--) {
                        double grdu = grdl;
                        grdl = (B(i + 1, k) - B(i, k)) * dxm1;
                        double buoyu = buoyl;
                        buoyl = 0.5 * (B(i + 1, k) + B(i, k));
                        double dxz = (zz[IJK(i + 1, j, k)] - zz[IJK(i, j, k)]) * dxm1;
                        prgr = prgr + 0.5 * (grdu + grdl) * 0.5 *
                                          (hun[IJK(i, j, k)] + hun[IJK(i, j, k + 1)]) -
                               dxz * (buoyu - buoyl);
                        idpdx[IJK(i, j, k)] = hun[IJK(i, j, k)] * prgr;
                    }
#undef B
                }
#pragma omp for schedule(static)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++)
                if (av[IJ(i, j)] >= 1) {
#define B(jj, k) buoy[IJK(i, jj, k)]
                    double grdl = (B(j + 1, kmax) - B(j, kmax)) * dym1;
                    double buoyl = 0.5 * (B(j + 1, kmax) + B(j, kmax));
                    double prgr = grdl * 0.5 * hvn[IJK(i, j, kmax)];
                    idpdy[IJK(i, j, kmax)] = hvn[IJK(i, j, kmax)] * prgr;
                    for (int k = kmax - 1; k >= 1; k--) {
                        double grdu = grdl;
                        grdl = (B(j + 1, k) - B(j, k)) * dym1;
                        double buoyu = buoyl;
                        buoyl = 0.5 * (B(j + 1, k) + B(j, k));
                        double dyz = (zz[IJK(i, j + 1, k)] - zz[IJK(i, j, k)]) * dym1;
                        prgr = prgr + 0.5 * (grdu + grdl) * 0.5 *
                                          (hvn[IJK(i, j, k)] + hvn[IJK(i, j, k + 1)]) -
                               dyz * (buoyu - buoyl);
                        idpdy[IJK(i, j, k)] = hvn[IJK(i, j, k)] * prgr;
                    }
#undef B
                }
    }
}

/* ---- a10: ip_shchepetkin_mcwilliams.F90:71-288 --------------------------
 * eps, OneFifth and 1./12. are single-precision literals in the reference
 * (promoted to FP64), so their FP32-rounded values are used here. */
void or_ip_shchepetkin_mcwilliams(const or_dom *d, const int32_t *az,
                                  const int32_t *au, const int32_t *av,
                                  const double *H, const double *hn,
                                  const double *hun, const double *hvn,
                                  const double *buoy, const double *sseo,
                                  double *zz, double *idpdx, double *idpdy)
{
    DOMVARS(d);
    const double eps = (double)1.e-10f, OneFifth = (double)0.2f;
    const double c12 = (double)(1.0f / 12.0f);
    const double dxm1 = 1.0 / d->dx, dym1 = 1.0 / d->dy;
    const long N3 = NXY * (kmax + 1);
    double *dR = scr(SCR_DR, (size_t)N3), *dZ = scr(SCR_DZ, (size_t)N3);
    double *P = scr(SCR_P, (size_t)N3);
    double *dZx = scr(SCR_DZX, (size_t)NXY), *dRx = scr(SCR_DRX, (size_t)NXY);

    for (long c = 0; c < NXY; c++) {
        zz[c] = 0.0;
        idpdx[c] = 0.0;
        idpdy[c] = 0.0;
    }
#define Z(i, j, k) zz[IJK(i, j, k)]
#define BU(i, j, k) buoy[IJK(i, j, k)]
#pragma omp parallel OMP_NT
    {
#pragma omp for schedule(static)
        for (int j = jl - HALO; j <= jh + HALO; j++)
            for (int i = il - HALO; i <= ih + HALO; i++)
                if (az[IJ(i, j)] >= 1) {
                    Z(i, j, 1) = -H[IJ(i, j)] + 0.5 * hn[IJK(i, j, 1)];
                    for (int k = 2; k <= kmax; k++)
                        Z(i, j, k) = Z(i, j, k - 1) +
                                     0.5 * (hn[IJK(i, j, k - 1)] + hn[IJK(i, j, k)]);
                }
#pragma omp for schedule(static)
        for (int j = jl; j <= jh + 1; j++)
            for (int i = il; i <= ih + 1; i++)
                if (az[IJ(i, j)] >= 1) {
#define R_(k) dR[IJK(i, j, k)]
#define Z_(k) dZ[IJK(i, j, k)]
                    for (int k = 1; k <= kmax - 1; k++) {
                        R_(k) = BU(i, j, k + 1) - BU(i, j, k);
                        Z_(k) = Z(i, j, k + 1) - Z(i, j, k);
                    }
                    R_(kmax) = R_(kmax - 1);
                    Z_(kmax) = Z_(kmax - 1);
                    R_(0) = R_(1);
                    Z_(0) = Z_(1);
                    for (int k = kmax; k >= 1; k--) {
                        double cff = 2.0 * R_(k) * R_(k - 1);
                        if (cff > eps) R_(k) = cff / (R_(k) + R_(k - 1));
                        else           R_(k) = 0.0;
                        Z_(k) = 2.0 * Z_(k) * Z_(k - 1) / (Z_(k) + Z_(k - 1));
                    }
                    double cff1 = 1.0 / (Z(i, j, kmax) - Z(i, j, kmax - 1));
                    double cff2 = 0.5 * (BU(i, j, kmax) - BU(i, j, kmax - 1)) * 0.5 *
                                  hn[IJK(i, j, kmax)] * cff1;
                    P[IJK(i, j, kmax)] = (BU(i, j, kmax) + cff2) * 0.5 * hn[IJK(i, j, kmax)];
                    for (int k = kmax - 1; k >= 1; k--)
                        P[IJK(i, j, k)] =
                            P[IJK(i, j, k + 1)] +
                            0.5 * ((BU(i, j, k + 1) + BU(i, j, k)) * (Z(i, j, k + 1) - Z(i, j, k)) -
                                   OneFifth * ((R_(k + 1) - R_(k)) *
                                                   (Z(i, j, k + 1) - Z(i, j, k) - c12 * (Z_(k + 1) + Z_(k))) -
                                               (Z_(k + 1) - Z_(k)) *
                                                   (BU(i, j, k + 1) - BU(i, j, k) - c12 * (R_(k + 1) + R_(k)))));
#undef R_
#undef Z_
                }
        /* x-direction, one k at a time (ip_shchepetkin_mcwilliams.F90:153-212) */
        for (int k = kmax; k >= 1; k--) {
#pragma omp for schedule(static)
            for (int j = jl; j <= jh; j++)
                for (int i = il - HALO + 1; i <= ih + HALO; i++) {
                    dZx[IJ(i, j)] = Z(i, j, k) - Z(i - 1, j, k);
                    dRx[IJ(i, j)] = BU(i, j, k) - BU(i - 1, j, k);
                }
#pragma omp for schedule(static)
            for (int j = jl; j <= jh; j++)
                for (int i = il; i <= ih + HALO - 1; i++) {
                    double cff = 2.0 * dZx[IJ(i, j)] * dZx[IJ(i + 1, j)];
                    if (cff > eps) {
                        double cff1 = 1.0 / (dZx[IJ(i, j)] + dZx[IJ(i + 1, j)]);
                        dZx[IJ(i, j)] = cff * cff1;
                    } else {
                        dZx[IJ(i, j)] = 0.0;
                    }
                    double cff1 = 2.0 * dRx[IJ(i, j)] * dRx[IJ(i + 1, j)];
                    if (cff1 > eps) {
                        double cff2 = 1.0 / (dRx[IJ(i, j)] + dRx[IJ(i + 1, j)]);
                        dRx[IJ(i, j)] = cff1 * cff2;
                    } else {
                        dRx[IJ(i, j)] = 0.0;
                    }
                }
#pragma omp for schedule(static)
            for (int j = jl; j <= jh; j++)
                for (int i = il; i <= ih + HALO - 2; i++)
                    if (au[IJ(i, j)] >= 1) {
                        double AJ = P[IJK(i + 1, j, k)] - P[IJK(i, j, k)];
                        double FC =
                            0.5 * ((BU(i + 1, j, k) + BU(i, j, k)) * (Z(i + 1, j, k) - Z(i, j, k)) -
                                   OneFifth * ((dRx[IJ(i + 1, j)] - dRx[IJ(i, j)]) *
                                                   (Z(i + 1, j, k) - Z(i, j, k) -
                                                    c12 * (dZx[IJ(i + 1, j)] + dZx[IJ(i, j)])) -
                                               (dZx[IJ(i + 1, j)] - dZx[IJ(i, j)]) *
                                                   (BU(i + 1, j, k) - BU(i, j, k) -
                                                    c12 * (dRx[IJ(i + 1, j)] + dRx[IJ(i, j)]))));
                        AJ = AJ + FC;
                        double h = hun[IJK(i, j, k)];
                        idpdx[IJK(i, j, k)] =
                            AJ * dxm1 * h - h * dxm1 * (sseo[IJ(i + 1, j)] - sseo[IJ(i, j)]) * 0.5 *
                                                (BU(i + 1, j, kmax) + BU(i, j, kmax));
                    }
        }
        /* y-direction (ip_shchepetkin_mcwilliams.F90:214-287) */
        for (int k = kmax; k >= 1; k--) {
#pragma omp for schedule(static)
            for (int j = jl - HALO + 1; j <= jh + HALO; j++)
                for (int i = il; i <= ih; i++) {
                    dZx[IJ(i, j)] = Z(i, j, k) - Z(i, j - 1, k);
                    dRx[IJ(i, j)] = BU(i, j, k) - BU(i, j - 1, k);
                }
            /* the in-place harmonic mean reads row j+1 before it is
             * overwritten, so the rows are processed in ascending order by
             * one thread each column strip */
#pragma omp for schedule(static)
            for (int i = il; i <= ih; i++)
                for (int j = jl; j <= jh + HALO - 1; j++) {
                    double cff = 2.0 * dZx[IJ(i, j)] * dZx[IJ(i, j + 1)];
                    if (cff > eps) {
                        double cff1 = 1.0 / (dZx[IJ(i, j)] + dZx[IJ(i, j + 1)]);
                        dZx[IJ(i, j)] = cff * cff1;
                    } else {
                        dZx[IJ(i, j)] = 0.0;
                    }
                    double cff1 = 2.0 * dRx[IJ(i, j)] * dRx[IJ(i, j + 1)];
                    if (cff1 > eps) {
                        double cff2 = 1.0 / (dRx[IJ(i, j)] + dRx[IJ(i, j + 1)]);
                        dRx[IJ(i, j)] = cff1 * cff2;
                    } else {
                        dRx[IJ(i, j)] = 0.0;
                    }
                }
#pragma omp for schedule(static)
            for (int j = jl; j <= jh + HALO - 2; j++)
                for (int i = il; i <= ih; i++)
                    if (av[IJ(i, j)] >= 1) {
                        double AJ = P[IJK(i, j + 1, k)] - P[IJK(i, j, k)];
                        double FC =
                            0.5 * ((BU(i, j + 1, k) + BU(i, j, k)) * (Z(i, j + 1, k) - Z(i, j, k)) -
                                   OneFifth * ((dRx[IJ(i, j + 1)] - dRx[IJ(i, j)]) *
                                                   (Z(i, j + 1, k) - Z(i, j, k) -
                                                    c12 * (dZx[IJ(i, j + 1)] + dZx[IJ(i, j)])) -
                                               (dZx[IJ(i, j + 1)] - dZx[IJ(i, j)]) *
                                                   (BU(i, j + 1, k) - BU(i, j, k) -
                                                    c12 * (dRx[IJ(i, j + 1)] + dRx[IJ(i, j)]))));
                        AJ = AJ + FC;
                        double h = hvn[IJK(i, j, k)];
                        idpdy[IJK(i, j, k)] =
                            AJ * dym1 * h - h * dym1 * (sseo[IJ(i, j + 1)] - sseo[IJ(i, j)]) * 0.5 *
                                                (BU(i, j + 1, kmax) + BU(i, j, kmax));
                    }
        }
    }
#undef Z
#undef BU
}

/* ---- a11: the SUBSTR_INI_PRESS part of do_internal_pressure,
 * internal_pressure.F90:235-258.  *first is the routine's SAVE'd flag. */
void or_substr_ini_press(const or_dom *d, int *first, double *idpdx,
                         double *idpdy, double *idpdx0, double *idpdy0)
{
    DOMVARS(d);
    for (int k = 0; k <= kmax; k++)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                long c = IJK(i, j, k);
                if (*first) {
                    idpdx0[c] = idpdx[c];
                    idpdx[c] = 0.0;
                    idpdy0[c] = idpdy[c];
                    idpdy[c] = 0.0;
                } else {
                    idpdx[c] = idpdx[c] - idpdx0[c];
                    idpdy[c] = idpdy[c] - idpdy0[c];
                }
            }
    *first = 0;
}

/* ---- a13: halo_zones.F90:81-187, serial ------------------------------- */
void or_update_3d_halo(const or_dom *d, double *f1, const double *f2,
                       int mirror)
{
    DOMVARS(d);
    if (mirror != 1) return;
    /* four whole-array assignments, in order; each right-hand side is read
     * in full before its assignment (Fortran array semantics) */
    for (int k = 0; k <= kmax; k++) {
        for (int j = jl - HALO; j <= jh + HALO; j++)
            f1[IJK(il - 1, j, k)] = f2[IJK(il, j, k)];
        for (int j = jl - HALO; j <= jh + HALO; j++)
            f1[IJK(ih + 1, j, k)] = f2[IJK(ih, j, k)];
        for (int i = il - HALO; i <= ih + HALO; i++)
            f1[IJK(i, jl - 1, k)] = f2[IJK(i, jl, k)];
        for (int i = il - HALO; i <= ih + HALO; i++)
            f1[IJK(i, jh + 1, k)] = f2[IJK(i, jh, k)];
    }
}

void or_update_2d_halo(const or_dom *d, double *f1, const double *f2,
                       int mirror, int *do_mirror_state)
{
    DOMVARS(d);
    if (mirror >= 0) *do_mirror_state = mirror != 0;
    if (!*do_mirror_state) return;
    for (int j = jl - HALO; j <= jh + HALO; j++) f1[IJ(il - 1, j)] = f2[IJ(il, j)];
    for (int j = jl - HALO; j <= jh + HALO; j++) f1[IJ(ih + 1, j)] = f2[IJ(ih, j)];
    for (int i = il - HALO; i <= ih + HALO; i++) f1[IJ(i, jl - 1)] = f2[IJ(i, jl)];
    for (int i = il - HALO; i <= ih + HALO; i++) f1[IJ(i, jh + 1)] = f2[IJ(i, jh)];
    f1[IJ(il - 1, jh + 1)] = f2[IJ(il, jh)];
    f1[IJ(ih + 1, jh + 1)] = f2[IJ(ih, jh)];
    f1[IJ(ih + 1, jl - 1)] = f2[IJ(ih, jl)];
    f1[IJ(il - 1, jl - 1)] = f2[IJ(il, jl)];
}

void or_decomp_1d(int n, int np, int m, int *s, int *e)
{
    int nlocal = n / np;
    *s = m * nlocal + 1;
    int deficit = n % np;
    *s = *s + (m < deficit ? m : deficit);
    if (m < deficit) nlocal = nlocal + 1;
    *e = *s + nlocal - 1;
    if (*e > n || m == np - 1) *e = n;
}

void or_exchange_tiles(int px, int py, const int *ni, const int *nj, int nk,
                       double **tiles)
{
    /* receive-side view: every ghost cell of tile (tx,ty) that lies in a
     * neighbour's interior takes that neighbour's value; tiles in one row
     * share nj, tiles in one column share ni */
    for (int ty = 0; ty < py; ty++)
        for (int tx = 0; tx < px; tx++) {
            int t = ty * px + tx;
            long NX = ni[t] + 2 * HALO, NY = nj[t] + 2 * HALO;
            for (int k = 0; k < nk; k++)
                for (long jj = 0; jj < NY; jj++)
                    for (long ii = 0; ii < NX; ii++) {
                        int ox = ii < HALO ? -1 : (ii >= ni[t] + HALO ? 1 : 0);
                        int oy = jj < HALO ? -1 : (jj >= nj[t] + HALO ? 1 : 0);
                        if (!ox && !oy) continue;
                        int sx = tx + ox, sy = ty + oy;
                        if (sx < 0 || sx >= px || sy < 0 || sy >= py) continue;
                        int s = sy * px + sx;
                        long SNX = ni[s] + 2 * HALO, SNY = nj[s] + 2 * HALO;
                        long si = ox < 0 ? ii + ni[s] : (ox > 0 ? ii - ni[t] : ii);
                        long sj = oy < 0 ? jj + nj[s] : (oy > 0 ? jj - nj[t] : jj);
                        tiles[t][ii + NX * (jj + NY * k)] =
                            tiles[s][si + SNX * (sj + SNY * k)];
                    }
        }
}

/* ---- a15: budgets (error-free double-double accumulation) --------------- */
static inline void two_sum(double a, double b, double *s, double *e)
{
    double x = a + b, bv = x - a, av = x - bv;
    *s = x;
    *e = (a - av) + (b - bv);
}
static inline void two_prod(double a, double b, double *p, double *e)
{
    *p = a * b;
    *e = fma(a, b, -*p);
}
typedef struct { double hi, lo; } dd;
static inline dd dd_add_d(dd a, double b)
{
    double s, e;
    two_sum(a.hi, b, &s, &e);
    e += a.lo;
    dd r;
    two_sum(s, e, &r.hi, &r.lo);
    return r;
}

void or_budget(const or_dom *d, const int32_t *az, const double *h,
               const double *f, double *mass, double *volume)
{
    DOMVARS(d);
    const double area = d->dx * d->dy;
    dd m = {0, 0}, v = {0, 0};
    for (int k = 1; k <= kmax; k++)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                if (az[IJ(i, j)] != 1) continue;
                long c = IJK(i, j, k);
                double p, pe;
                two_prod(h[c], f ? f[c] : 1.0, &p, &pe);
                m = dd_add_d(m, p);
                m = dd_add_d(m, pe);
                v = dd_add_d(v, h[c]);
            }
    /* scale once: the area factor is common to every term */
    double mh, ml, vh, vl;
    two_prod(m.hi, area, &mh, &ml);
    ml += m.lo * area;
    two_prod(v.hi, area, &vh, &vl);
    vl += v.lo * area;
    *mass = mh + ml;
    *volume = vh + vl;
}

/* ---- input producers of the transport step --------------------------------
 * start_macro (elevations, depths and mean transports of the macro step),
 * coordinates (layer heights on T, U and V points) and ww_momentum_3d (the
 * grid-related vertical velocity), which m3d.F90:446-548 runs before the
 * tracer advection.  Cartesian metrics; kmin = 1 (sigma and general
 * coordinates). */

static inline double dmax2(double a, double b) { return a > b ? a : b; }
static inline double dmin2(double a, double b) { return a < b ? a : b; }

/* start_macro, start_macro.F90:54-86 */
void or_start_macro(const or_dom *d, int M, double min_depth, const double *H,
                    const double *HU, const double *HV, const double *z,
                    double *sseo, double *ssen, double *ssuo, double *ssun,
                    double *ssvo, double *ssvn, double *Dn, double *Dun,
                    double *Dvn, double *Uint, double *Vint)
{
    DOMVARS(d);
    for (long c = 0; c < NXY; c++) {
        sseo[c] = ssen[c];
        ssen[c] = z[c];
        Dn[c] = ssen[c] + H[c];
    }
    for (long jj = 0; jj < NY; jj++)
        for (long ii = 0; ii < NX - 1; ii++) {
            const long c = ii + NX * jj;
            ssuo[c] = ssun[c];
            ssun[c] = 0.25 * (sseo[c] + sseo[c + 1] + ssen[c] + ssen[c + 1]);
            ssun[c] = dmax2(ssun[c], -HU[c] + min_depth);
            Dun[c] = ssun[c] + HU[c];
        }
    for (long jj = 0; jj < NY - 1; jj++)
        for (long ii = 0; ii < NX; ii++) {
            const long c = ii + NX * jj;
            ssvo[c] = ssvn[c];
            ssvn[c] = 0.25 * (sseo[c] + sseo[c + NX] + ssen[c] + ssen[c + NX]);
            ssvn[c] = dmax2(ssvn[c], -HV[c] + min_depth);
            Dvn[c] = ssvn[c] + HV[c];
        }
    const double split = 1.0 / (double)(float)M;
    for (long c = 0; c < NXY; c++) {
        Uint[c] = split * Uint[c];
        Vint[c] = split * Vint[c];
    }
}

/* the SAVE'd state of coordinates(), sigma_coordinates() and
 * general_coordinates(): the first-call flag and the level arrays */
typedef struct {
    int vert_cord, gamma_surf, first, equiv_sigma, kmax;
    double ddu, ddl, d_gamma, min_depth;
    double *ga, *dga, *sig, *be, *gga;
} or_coord;

void *or_coord_new(const or_dom *d, int vert_cord, double ddu, double ddl,
                   double d_gamma, int gamma_surf, double min_depth)
{
    DOMVARS(d);
    or_coord *s = (or_coord *)calloc(1, sizeof(or_coord));
    s->vert_cord = vert_cord;
    s->ddu = ddu;
    s->ddl = ddl;
    s->d_gamma = d_gamma;
    s->gamma_surf = gamma_surf;
    s->min_depth = min_depth;
    s->first = 1;
    s->kmax = kmax;
    s->ga = (double *)calloc(4 * (size_t)(kmax + 1), sizeof(double));
    s->dga = s->ga + (kmax + 1);
    s->sig = s->dga + (kmax + 1);
    s->be = s->sig + (kmax + 1);
    if (vert_cord == 3)
        s->gga = (double *)calloc((size_t)(NXY * (kmax + 1)), sizeof(double));
    return s;
}

void or_coord_delete(void *p)
{
    or_coord *s = (or_coord *)p;
    if (!s) return;
    free(s->ga);
    free(s->gga);
    free(s);
}

/* the surface/bottom zooming of Garapon, sigma_coordinates.F90:72-76 and
 * general_coordinates.F90:72-75 */
static double zoom_level(double ddu, double ddl, int k, int kmax)
{
    double g = tanh((ddl + ddu) * k / (double)(float)kmax - ddl) + tanh(ddl);
    return g / (tanh(ddl) + tanh(ddu)) - 1.0;
}

/* coordinates(), coordinates.F90:66-107, for vert_cord 1 (sigma,
 * sigma_coordinates.F90:53-147), 2 (z-levels: nothing but zc) and 3
 * (general, general_coordinates.F90:56-210).  zc may be NULL.  Returns 0,
 * 1 when the general coordinates are not monotone (the reference stops,
 * :92-99), 2 for another vert_cord. */
int or_coordinates(const or_dom *d, void *state, double dt, double cord_relax,
                   double maxdepth, const double *H, const double *HU,
                   const double *HV, const double *sseo, const double *ssen,
                   const double *ssuo, const double *ssun, const double *ssvo,
                   const double *ssvn, double *ho, double *hn, double *huo,
                   double *hun, double *hvo, double *hvn, double *zc)
{
    DOMVARS(d);
    or_coord *s = (or_coord *)state;
    double *ga = s->ga, *dga = s->dga, *sig = s->sig, *be = s->be,
           *gga = s->gga;
#define G3(c, k) gga[(c) + NXY * (long)(k)]
#define F3(a, c, k) a[(c) + NXY * (long)(k)]
    if (s->vert_cord == 1) {
        if (s->first) {
            if (s->ddu <= 0.0 && s->ddl <= 0.0) {
                s->equiv_sigma = 1;
                ga[0] = -1.0;
                for (int k = 1; k <= kmax; k++) ga[k] = ga[k - 1] + 1.0 / kmax;
                ga[kmax] = 0.0;
            } else {
                if (s->ddu < 0.0) s->ddu = 0.0;
                if (s->ddl < 0.0) s->ddl = 0.0;
                ga[0] = -1.0;
                dga[0] = 0.0;
                for (int k = 1; k <= kmax; k++) {
                    ga[k] = zoom_level(s->ddu, s->ddl, k, kmax);
                    dga[k] = ga[k] - ga[k - 1];
                }
            }
            s->first = 0;
        }
        if (s->equiv_sigma) {
            const double kmaxm1 = 1.0 / (double)(float)kmax;
            for (long c = 0; c < NXY; c++)
                for (int k = 0; k <= kmax; k++) {
                    F3(ho, c, k) = (sseo[c] + H[c]) * kmaxm1;
                    F3(hn, c, k) = (ssen[c] + H[c]) * kmaxm1;
                }
            for (long jj = 0; jj < NY; jj++)
                for (long ii = 0; ii < NX - 1; ii++) {
                    const long c = ii + NX * jj;
                    for (int k = 0; k <= kmax; k++) {
                        F3(huo, c, k) = (ssuo[c] + HU[c]) * kmaxm1;
                        F3(hun, c, k) = (ssun[c] + HU[c]) * kmaxm1;
                    }
                }
            for (long jj = 0; jj < NY - 1; jj++)
                for (long ii = 0; ii < NX; ii++) {
                    const long c = ii + NX * jj;
                    for (int k = 0; k <= kmax; k++) {
                        F3(hvo, c, k) = (ssvo[c] + HV[c]) * kmaxm1;
                        F3(hvn, c, k) = (ssvn[c] + HV[c]) * kmaxm1;
                    }
                }
        } else {
            for (int k = 1; k <= kmax; k++) {
                for (long c = 0; c < NXY; c++) {
                    F3(ho, c, k) = (sseo[c] + H[c]) * dga[k];
                    F3(hn, c, k) = (ssen[c] + H[c]) * dga[k];
                }
                for (long jj = 0; jj < NY; jj++)
                    for (long ii = 0; ii < NX - 1; ii++) {
                        const long c = ii + NX * jj;
                        F3(huo, c, k) = (ssuo[c] + HU[c]) * dga[k];
                        F3(hun, c, k) = (ssun[c] + HU[c]) * dga[k];
                    }
                for (long jj = 0; jj < NY - 1; jj++)
                    for (long ii = 0; ii < NX; ii++) {
                        const long c = ii + NX * jj;
                        F3(hvo, c, k) = (ssvo[c] + HV[c]) * dga[k];
                        F3(hvn, c, k) = (ssvn[c] + HV[c]) * dga[k];
                    }
            }
        }
    } else if (s->vert_cord == 3) {
        if (s->first) {
            for (int k = 0; k <= kmax; k++) ga[k] = k;
            be[0] = -1.0;
            sig[0] = -1.0;
            if (s->ddu < 0.0) s->ddu = 0.0;
            if (s->ddl < 0.0) s->ddl = 0.0;
            for (int k = 1; k <= kmax; k++) {
                be[k] = zoom_level(s->ddu, s->ddl, k, kmax);
                /* k/float(kmax) is a default-REAL (FP32) quotient */
                sig[k] = (double)((float)k / (float)kmax) - 1.0;
            }
            const int kk = s->gamma_surf ? kmax : 1;
            for (long c = 0; c < NXY; c++) {
                const double HH = dmax2(sseo[c] + H[c], s->min_depth);
                const double alpha =
                    dmin2(((be[kk] - be[kk - 1]) -
                           s->d_gamma / HH * (sig[kk] - sig[kk - 1])) /
                              ((be[kk] - be[kk - 1]) - (sig[kk] - sig[kk - 1])),
                          1.0);
                G3(c, 0) = -1.0;
                for (int k = 1; k <= kmax; k++) {
                    G3(c, k) = alpha * sig[k] + (1.0 - alpha) * be[k];
                    if (G3(c, k) < G3(c, k - 1)) return 1;
                }
            }
            for (int k = 1; k <= kmax; k++)
                for (long c = 0; c < NXY; c++) {
                    const double HH = dmax2(sseo[c] + H[c], s->min_depth);
                    F3(hn, c, k) = HH * (G3(c, k) - G3(c, k - 1));
                }
            for (int k = 1; k <= kmax; k++)
                for (long jj = 0; jj < NY; jj++)
                    for (long ii = 0; ii < NX - 1; ii++) {
                        const long c = ii + NX * jj;
                        const double HH = dmax2(ssuo[c] + HU[c], s->min_depth);
                        F3(huo, c, k) = HH * 0.5 *
                                        (G3(c, k) - G3(c, k - 1) + G3(c + 1, k) -
                                         G3(c + 1, k - 1));
                        F3(hun, c, k) = F3(huo, c, k);
                    }
            for (int k = 1; k <= kmax; k++)
                for (long jj = 0; jj < NY - 1; jj++)
                    for (long ii = 0; ii < NX; ii++) {
                        const long c = ii + NX * jj;
                        const double HH = dmax2(ssvo[c] + HV[c], s->min_depth);
                        F3(hvo, c, k) = HH * 0.5 *
                                        (G3(c, k) - G3(c, k - 1) + G3(c + NX, k) -
                                         G3(c + NX, k - 1));
                        F3(hvn, c, k) = F3(hvo, c, k);
                    }
            s->first = 0;
        }
        /* relaxation towards the target layer heights, :146-210 */
        for (long c = 0; c < NXY; c++) {
            const double r = cord_relax / dt * H[c] / maxdepth;
            const double HH = ssen[c] + H[c];
            if (HH < s->d_gamma) {
                for (int k = 1; k <= kmax; k++) {
                    F3(ho, c, k) = F3(hn, c, k);
                    F3(hn, c, k) = HH / kmax;
                }
            } else {
                double zz = -H[c];
                for (int k = 1; k <= kmax - 1; k++) {
                    F3(ho, c, k) = F3(hn, c, k);
                    F3(hn, c, k) = (F3(ho, c, k) * r + HH * (G3(c, k) - G3(c, k - 1))) /
                                   (r + 1.0);
                    zz = zz + F3(hn, c, k);
                }
                F3(ho, c, kmax) = F3(hn, c, kmax);
                F3(hn, c, kmax) = ssen[c] - zz;
            }
        }
        for (long jj = 0; jj < NY; jj++)
            for (long ii = 0; ii < NX - 1; ii++) {
                const long c = ii + NX * jj;
                const double r = cord_relax / dt * HU[c] / maxdepth;
                double zz = -HU[c];
                const double HH = ssun[c] + HU[c];
                for (int k = 1; k <= kmax - 1; k++) {
                    F3(huo, c, k) = F3(hun, c, k);
                    F3(hun, c, k) = (F3(huo, c, k) * r +
                                     HH * 0.5 *
                                         (G3(c, k) - G3(c, k - 1) + G3(c + 1, k) -
                                          G3(c + 1, k - 1))) /
                                    (r + 1.0);
                    zz = zz + F3(hun, c, k);
                }
                F3(huo, c, kmax) = F3(hun, c, kmax);
                F3(hun, c, kmax) = ssun[c] - zz;
            }
        for (long jj = 0; jj < NY - 1; jj++)
            for (long ii = 0; ii < NX; ii++) {
                const long c = ii + NX * jj;
                const double r = cord_relax / dt * HV[c] / maxdepth;
                double zz = -HV[c];
                const double HH = ssvn[c] + HV[c];
                for (int k = 1; k <= kmax - 1; k++) {
                    F3(hvo, c, k) = F3(hvn, c, k);
                    F3(hvn, c, k) = (F3(hvo, c, k) * r +
                                     HH * 0.5 *
                                         (G3(c, k) - G3(c, k - 1) + G3(c + NX, k) -
                                          G3(c + NX, k - 1))) /
                                    (r + 1.0);
                    zz = zz + F3(hvn, c, k);
                }
                F3(hvo, c, kmax) = F3(hvn, c, kmax);
                F3(hvn, c, kmax) = ssvn[c] - zz;
            }
    } else if (s->vert_cord != 2) {
        return 2;
    }
    s->first = 0;
    if (zc) {
        for (long c = 0; c < NXY; c++) {
            F3(zc, c, 0) = -H[c];
            F3(zc, c, 1) = -H[c] + 0.5 * F3(hn, c, 1);
            for (int k = 2; k <= kmax; k++)
                F3(zc, c, k) = F3(zc, c, k - 1) + 0.5 * (F3(hn, c, k - 1) + F3(hn, c, k));
        }
    }
#undef G3
#undef F3
    return 0;
}

/* ww_momentum_3d, ww_momentum_3d.F90:66-119 (Cartesian, kmin = 1): the
 * grid-related vertical velocity from the layer continuity equation, upward
 * from ww(0) (kept), k = 1..kmax-1 on interior water columns.  The serial
 * 3-D halo update that follows is a no-op. */
void or_ww_momentum_3d(const or_dom *d, const int32_t *az, double dt,
                       const double *uu, const double *vv, const double *ho,
                       const double *hn, double *ww)
{
    DOMVARS(d);
    const double dtm1 = 1.0 / dt, dx = d->dx, dy = d->dy, ard1 = d->ard1;
    for (int k = 1; k <= kmax - 1; k++)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                if (az[IJ(i, j)] != 1) continue;
                const long c = IJK(i, j, k);
                ww[c] = ww[c - NXY] - (hn[c] - ho[c]) * dtm1 -
                        (uu[c] * dy - uu[IJK(i - 1, j, k)] * dy + vv[c] * dx -
                         vv[IJK(i, j - 1, k)] * dx) *
                            ard1;
            }
}

/* ---- the step after advection: vertical diffusion ------------------------ */

/* getm_tridiagonal, src/futils/tridiagonal.F90:8-51: rows fi..lt of a
 * tridiagonal system, au sub-, bu main, cu super-diagonal, du right-hand
 * side; elimination from lt down to fi, then substitution upwards.  ru/qu
 * are work arrays of kmax+1. */
void or_tridiagonal(int kmax, int fi, int lt, const double *au, const double *bu,
                    const double *cu, const double *du, double *value)
{
    double *ru = (double *)malloc(2 * (size_t)(kmax + 1) * sizeof(double));
    double *qu = ru + (kmax + 1);
    ru[lt] = au[lt] / bu[lt];
    qu[lt] = du[lt] / bu[lt];
    for (int i = lt - 1; i >= fi + 1; i--) {
        ru[i] = au[i] / (bu[i] - cu[i] * ru[i + 1]);
        qu[i] = (du[i] - cu[i] * qu[i + 1]) / (bu[i] - cu[i] * ru[i + 1]);
    }
    qu[fi] = (du[fi] - cu[fi] * qu[fi + 1]) / (bu[fi] - cu[fi] * ru[fi + 1]);
    value[fi] = qu[fi];
    for (int i = fi + 1; i <= lt; i++) value[i] = qu[i] - ru[i] * value[i - 1];
    free(ru);
}

/* the implicit vertical diffusion of one column (temperature.F90:505-547,
 * salinity.F90:351-387); rad1d = NULL for salinity (no source terms) */
static void diffuse_column(int kmax, double dt, double cnpar, double avm,
                           const double *hn, const double *nuh, long c,
                           long NXY, double *f, const double *rad1d,
                           double shf_dz, double dry_z, double *w)
{
#define C3(a, k) a[c + NXY * (long)(k)]
    double *auxo = w, *auxn = w + (kmax + 1), *a1 = w + 2 * (kmax + 1),
           *a2 = w + 3 * (kmax + 1), *a3 = w + 4 * (kmax + 1),
           *a4 = w + 5 * (kmax + 1), *res = w + 6 * (kmax + 1);
    for (int k = 1; k <= kmax - 1; k++) {
        auxo[k] = 2.0 * (1 - cnpar) * dt * (C3(nuh, k) + avm) /
                  (C3(hn, k + 1) + C3(hn, k));
        auxn[k] = 2.0 * cnpar * dt * (C3(nuh, k) + avm) /
                  (C3(hn, k + 1) + C3(hn, k));
    }
    int k = kmax;
    a1[k] = -auxn[k - 1];
    a2[k] = C3(hn, k) + auxn[k - 1];
    a4[k] = C3(f, k) * (C3(hn, k) - auxo[k - 1]) + C3(f, k - 1) * auxo[k - 1];
    if (rad1d) a4[k] = a4[k] + dry_z * dt * (rad1d[k] + shf_dz - rad1d[k - 1]);
    for (k = 2; k <= kmax - 1; k++) {
        a3[k] = -auxn[k];
        a1[k] = -auxn[k - 1];
        a2[k] = C3(hn, k) + auxn[k] + auxn[k - 1];
        a4[k] = C3(f, k + 1) * auxo[k] +
                C3(f, k) * (C3(hn, k) - auxo[k] - auxo[k - 1]) +
                C3(f, k - 1) * auxo[k - 1];
        if (rad1d) a4[k] = a4[k] + dry_z * dt * (rad1d[k] - rad1d[k - 1]);
    }
    k = 1;
    a3[k] = -auxn[k];
    a2[k] = C3(hn, k) + auxn[k];
    a4[k] = C3(f, k + 1) * auxo[k] + C3(f, k) * (C3(hn, k) - auxo[k]);
    if (rad1d) a4[k] = a4[k] + dry_z * dt * (rad1d[k] - rad1d[k - 1]);
    or_tridiagonal(kmax, 1, kmax, a1, a2, a3, a4, res);
    for (k = 1; k <= kmax; k++) C3(f, k) = res[k];
#undef C3
}

/* the vertical diffusion of do_salinity, salinity.F90:346-393 (after the
 * advection): Crank-Nicolson (cnpar) with nuh + avmols on interior water
 * columns, kmax > 1 */
void or_salinity_diffusion(const or_dom *d, const int32_t *az, double dt,
                           double cnpar, const double *hn, const double *nuh,
                           double *S)
{
    DOMVARS(d);
    const double avmols = (double)1.1e-9f; /* parameters.F90:27 */
    if (kmax <= 1) return;
#pragma omp parallel OMP_NT
    {
        double *w = (double *)malloc(7 * (size_t)(kmax + 1) * sizeof(double));
#pragma omp for schedule(static)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++)
                if (az[IJ(i, j)] == 1)
                    diffuse_column(kmax, dt, cnpar, avmols, hn, nuh, IJ(i, j), NXY, S,
                                   NULL, 0.0, 0.0, w);
        free(w);
    }
}

/* the solar radiation and vertical diffusion of do_temperature,
 * temperature.F90:463-551 (after the advection): the short-wave profile
 * rad = swr (A exp(-z/g1) + (1-A) exp(-z/g2)) on columns with az >= 1, then
 * on water columns (az == 1) the primitive sea-ice clip of the surface heat
 * flux, the optional bottom reflection, and the implicit diffusion with the
 * radiation divergence and the surface flux as sources.  2-D inputs: swr,
 * shf, A, g1, g2, dry_z; S is read at kmax only.  rad is written. */
void or_temperature_diffusion(const or_dom *d, const int32_t *az, double dt,
                              double cnpar, const double *hn, const double *nuh,
                              const double *swr, const double *shf,
                              const double *A, const double *g1, const double *g2,
                              const double *dry_z, const double *S,
                              double swr_bot_refl_frac, double swr_min_bot_frac,
                              double *rad, double *T)
{
    DOMVARS(d);
    const double avmolt = (double)1.4e-7f; /* parameters.F90:26 */
    const double rho_0 = (double)1025.f, cp = (double)3985.f;
    const double rho_0_cpi = 1.0 / (rho_0 * cp);
#define C3(a, c, k) a[(c) + NXY * (long)(k)]
#pragma omp parallel for OMP_NT schedule(static)
    for (int j = jl; j <= jh; j++)
        for (int i = il; i <= ih; i++) {
            const long c = IJ(i, j);
            if (az[c] < 1) continue;
            const double swr_loc = swr[c];
            C3(rad, c, kmax) = swr_loc;
            double zz = 0.0;
            for (int k = kmax - 1; k >= 0; k--) {
                zz = zz + C3(hn, c, k + 1);
                C3(rad, c, k) = swr_loc * (A[c] * exp(-zz / g1[c]) +
                                           (1 - A[c]) * exp(-zz / g2[c]));
            }
        }
#pragma omp parallel OMP_NT
    {
        double *w = (double *)malloc(8 * (size_t)(kmax + 1) * sizeof(double));
        double *rad1d = w + 7 * (kmax + 1);
#pragma omp for schedule(static)
        for (int j = jl; j <= jh; j++)
            for (int i = il; i <= ih; i++) {
                const long c = IJ(i, j);
                if (az[c] != 1) continue;
                double shf_loc = shf[c];
                if (C3(T, c, kmax) <= -0.0575f * C3(S, c, kmax))
                    shf_loc = shf_loc < 0.0 ? 0.0 : shf_loc; /* max(0, shf) */
                for (int k = 0; k <= kmax; k++) rad1d[k] = C3(rad, c, k);
                if (swr_bot_refl_frac > 0.0 &&
                    rad1d[0] / rad1d[kmax] > swr_min_bot_frac) {
                    const double swr_refl = rad1d[0] * swr_bot_refl_frac;
                    rad1d[0] = rad1d[0] - swr_refl;
                    double zz = 0.0;
                    for (int k = 1; k <= kmax; k++) {
                        zz = zz + C3(hn, c, k);
                        rad1d[k] = C3(rad, c, k) -
                                   swr_refl * (A[c] * exp(-zz / g1[c]) +
                                               (1 - A[c]) * exp(-zz / g2[c]));
                    }
                }
                for (int k = 0; k <= kmax; k++) rad1d[k] = rad1d[k] * rho_0_cpi;
                if (kmax > 1)
                    diffuse_column(kmax, dt, cnpar, avmolt, hn, nuh, c, NXY, T, rad1d,
                                   shf_loc * rho_0_cpi, dry_z[c], w);
            }
        free(w);
    }
#undef C3
}

/* ---- a caller of the path: momentum advection ------------------------------
 * uv_advect_3d, src/3d/uv_advect_3d.F90:60-300 (all schemes but J7, no
 * numerical analyses): the velocity of the U (V) control volumes is
 * advected by do_advection_3d with the transports, vertical velocity and
 * layer heights averaged onto those volumes and tag U_TAG (V_TAG); the
 * advection terms are returned in uuEx (vvEx).  The reference's work arrays
 * are automatic, so the cells its loops leave unset hold whatever the stack
 * held; here they start at zero, and the V part sees what the U part left,
 * as in the reference. */
int or_uv_advect_3d(const or_dom *d, const or_advgrids *g, double dt,
                    const double *uu, const double *vv, const double *ww,
                    const double *ho, const double *hn, const double *hun,
                    const double *hvn, int split, int hscheme, int vscheme,
                    int adv_ver_iterations, double *uuEx, double *vvEx)
{
    DOMVARS(d);
    if (hscheme == OR_J7) return OR_EINVAL_SCHEME;
    if (hscheme == OR_NOADV && vscheme == OR_NOADV) return OR_OK;
    const long N3 = NXY * (kmax + 1);
    double *w = (double *)calloc(6 * (size_t)N3, sizeof(double));
    if (!w) return OR_ENOMEM;
    double *fadv = w, *uuadv = w + N3, *vvadv = w + 2 * N3, *wwadv = w + 3 * N3,
           *huadv = w + 4 * N3, *hvadv = w + 5 * N3;
    for (long c = 0; c < NXY; c++) wwadv[c] = 0.0;
    for (int k = 1; k <= kmax; k++)
        for (long jj = 0; jj < NY; jj++)
            for (long ii = 0; ii < NX - 1; ii++) {
                const long c = ii + NX * jj + NXY * k;
                fadv[c] = uu[c] / hun[c];
                uuadv[c] = 0.5 * (uu[c] + uu[c + 1]);
                vvadv[c] = 0.5 * (vv[c] + vv[c + 1]);
                wwadv[c] = 0.5 * (ww[c] + ww[c + 1]);
                huadv[c] = 0.5 * (ho[c + 1] + hn[c + 1]);
                hvadv[c] = 0.5 * (hvn[c] + hvn[c + 1]);
            }
    int rc = or_do_advection_3d(d, g, dt, fadv, uuadv, vvadv, wwadv, huadv, hvadv, hun, hun,
                                split, hscheme, vscheme, 0.0, OR_U_TAG,
                                adv_ver_iterations, NULL, uuEx, NULL);
    if (rc == OR_OK) {
        for (int k = 1; k <= kmax; k++)
            for (long jj = 0; jj < NY - 1; jj++)
                for (long ii = 0; ii < NX; ii++) {
                    const long c = ii + NX * jj + NXY * k;
                    fadv[c] = vv[c] / hvn[c];
                    uuadv[c] = 0.5 * (uu[c] + uu[c + NX]);
                    vvadv[c] = 0.5 * (vv[c] + vv[c + NX]);
                    wwadv[c] = 0.5 * (ww[c] + ww[c + NX]);
                    huadv[c] = 0.5 * (hun[c] + hun[c + NX]);
                    hvadv[c] = 0.5 * (ho[c + NX] + hn[c + NX]);
                }
        rc = or_do_advection_3d(d, g, dt, fadv, uuadv, vvadv, wwadv, huadv, hvadv, hvn, hvn,
                                split, hscheme, vscheme, 0.0, OR_V_TAG,
                                adv_ver_iterations, NULL, vvEx, NULL);
    }
    free(w);
    return rc;
}
