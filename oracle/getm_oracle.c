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
