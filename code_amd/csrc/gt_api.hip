// gt_api.hip -- the C-ABI of include/getm_transport.h.
//
// A context holds one domain: its masks on the device (the t_adv_grid sets
// of src/2d/advection.F90:263-318), the module array zz of
// src/3d/internal_pressure.F90, and scratch fields that stand in for the
// reference's automatic arrays (fi, hi, adv of do_advection_3d, the column
// work arrays of adv_split_w, dR/dZ/P of the SMcW scheme).  Scratch is
// allocated on first use and kept, so a time loop allocates nothing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/getm_transport.h"
#include "gt_adjoint.h"
#include "gt_substeps.h"
#include "gt_batched.h"
#include "gt_kernels.h"

using gt::Dims;

// blocks of the budgets' partial sums (the batched form uses the same number,
// so each of its outputs equals gt_budget_f32's bit for bit)
constexpr int kBudgetBlocks = 1024;

static thread_local char g_err[512] = "";

static int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return set_err(GT_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),     \
                     __FILE__, __LINE__);                                         \
  } while (0)

struct GridMasks {
  const uint8_t *uflux, *vflux, *uupd, *vupd, *fin;
  const int32_t* az;
  const uint8_t* flags;  // the five masks packed as gt::MB_* bits
};

struct gt_ctx {
  int imin, imax, jmin, jmax, kmax;
  Dims d;
  long N3;
  double dx, dy, ard1;
  hipStream_t st = nullptr;
  uint8_t* dmask = nullptr;
  int32_t* dint = nullptr;
  GridMasks H{}, U{}, V{};
  const int32_t *az = nullptr, *au = nullptr, *av = nullptr, *ax = nullptr;
  double *fi = nullptr, *hi = nullptr, *adv = nullptr, *tmp = nullptr;
  double *scr7 = nullptr, *zz = nullptr, *P = nullptr, *slice3 = nullptr;
  double *bpart = nullptr, *bout = nullptr;
  double* bbatch = nullptr;  // partials of gt_budget_batch_f32 (grown on demand)
  long bbatch_n = 0;
  double* noalloc = nullptr;  // this device's copy of the static scratch array
  unsigned long long* nwarn = nullptr;
  int adv_ver_iterations = 1;
  int do_mirror_2d = 1;
  int fused = 1;        // use the HALFSPLIT strip kernels where they apply
  int adj_keep = -1;    // reverse mode keeps each segment's sub-step states (-1: if they fit)
  int tile = 0;         // rows per strip of the two strip kernels (gt::strip_rows)
  int group = 3;        // tracers per wavefront of the FP32 batched kernels
  void* ws = nullptr;   // work fields of the strip kernels (grown on demand)
  size_t ws_bytes = 0;
  // coordinates(): namelist values (domain module) and the SAVE'd state
  int vert_cord = 0, gamma_surf = 1, coord_first = 1, equiv_sigma = 0;
  double ddu = 0.0, ddl = 0.0, d_gamma = 20.0, min_depth = 0.5;
  double* dlev = nullptr;  // device: dga (sigma) or sig | be (general), kmax+1 each
  double* gga = nullptr;   // device: gamma levels of the general coordinates, 3-D
  int* dbad = nullptr;     // device: set when the gamma levels fold over
  double* vdw = nullptr;   // vertical diffusion work: ru, qu, reflected profile (3 fields)
  double* tri = nullptr;   // batched tridiagonal work (grown on demand)
  long tri_n = 0;
  double* mwork = nullptr; // uv_advect_3d work: fadv uuadv vvadv wwadv huadv hvadv
  void* w32 = nullptr;     // per-sub-step and reverse-mode work (grown on demand)
  size_t w32_bytes = 0;
  // SUBSTR_INI_PRESS (internal_pressure.F90:132-141,235-258): the first
  // gradients (idpdx0 | idpdy0) and the SAVE'd first-call flag
  int sip = 0, sip_first = 1;
  double* idp0 = nullptr;
};

extern "C" const char* gt_last_error(void) { return g_err; }

// for the host-only translation units (gt_restart.cpp)
int gt_internal_set_error(int code, const char* msg) { return set_err(code, "%s", msg); }
extern "C" int gt_version(void) { return 1; }

static int ensure_ws(gt_ctx* c, size_t bytes) {
  if (c->ws_bytes >= bytes) return GT_OK;
  if (c->ws) {
    HIPCHK(hipDeviceSynchronize());  // the old fields may still be in use on any stream
    HIPCHK(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
  }
  HIPCHK(hipMalloc(&c->ws, bytes));
  c->ws_bytes = bytes;
  return GT_OK;
}

static int ensure(double*& p, long n) {
  if (p) return GT_OK;
  HIPCHK(hipMalloc(&p, sizeof(double) * (size_t)n));
  return GT_OK;
}
#define ENSURE(p, n)                 \
  do {                               \
    int r_ = ensure((p), (n));       \
    if (r_ != GT_OK) return r_;      \
  } while (0)

// ---- context ---------------------------------------------------------------
// domain.F90:294-346 on the host
static void masks_from_az(long NX, long NY, const int32_t* az, int32_t* au, int32_t* av,
                          int32_t* ax) {
  for (long n = 0; n < NX * NY; n++) au[n] = av[n] = ax[n] = 0;
  for (long j = 0; j < NY; j++)
    for (long i = 0; i < NX - 1; i++) {
      int a = az[i + NX * j], b = az[i + 1 + NX * j];
      int v = 0;
      if (a == 1 && b == 1) v = 1;
      if ((a == 1 && b == 2) || (a == 2 && b == 1)) v = 2;
      if (a == 2 && b == 2) v = 3;
      au[i + NX * j] = v;
    }
  for (long j = 0; j < NY - 1; j++)
    for (long i = 0; i < NX; i++) {
      int a = az[i + NX * j], b = az[i + NX * (j + 1)];
      int v = 0;
      if (a == 1 && b == 1) v = 1;
      if ((a == 1 && b == 2) || (a == 2 && b == 1)) v = 2;
      if (a == 2 && b == 2) v = 3;
      av[i + NX * j] = v;
    }
  for (long j = 0; j < NY - 1; j++)
    for (long i = 0; i < NX - 1; i++)
      if (az[i + NX * j] >= 1 && az[i + NX * (j + 1)] >= 1 && az[i + 1 + NX * j] >= 1 &&
          az[i + 1 + NX * (j + 1)] >= 1)
        ax[i + NX * j] = 1;
}

static int noalloc_scratch(double** p);

extern "C" int gt_create(gt_ctx** out, int imin, int imax, int jmin, int jmax, int kmax,
                         double dx, double dy, const int32_t* az, const int32_t* au,
                         const int32_t* av, const int32_t* ax) {
  if (!out || !az) return set_err(GT_EARG, "gt_create: NULL argument");
  *out = nullptr;
  if (imax < imin || jmax < jmin || kmax < 1 || !(dx > 0) || !(dy > 0))
    return set_err(GT_EARG, "gt_create: bad extents imin=%d imax=%d jmin=%d jmax=%d kmax=%d",
                   imin, imax, jmin, jmax, kmax);
  gt_ctx* c = new gt_ctx();
  c->imin = imin; c->imax = imax; c->jmin = jmin; c->jmax = jmax; c->kmax = kmax;
  c->d.NX = imax - imin + 1 + 2 * GT_HALO;
  c->d.NY = jmax - jmin + 1 + 2 * GT_HALO;
  c->d.kmax = kmax;
  c->d.NXY = (long)c->d.NX * c->d.NY;
  c->N3 = c->d.NXY * (kmax + 1);
  // default strip height: 16 rows, or 8 where 16-row strips would give the
  // strip kernels fewer than 4096 wavefronts (one tracer group), so a small
  // grid still fills the chip (gt_set_tuning / gt_autotune_batch override)
  {
    const long items16 = (long)((c->d.NX + 59) / 60) * ((c->d.NY + 15) / 16) * (kmax + 1);
    c->tile = items16 < 4096 ? 2 : 0;
  }
  c->dx = dx; c->dy = dy;
  c->ard1 = 1.0 / (dx * dy);  // domain.F90:596
  const long NX = c->d.NX, NY = c->d.NY, NXY = c->d.NXY;

  std::vector<int32_t> hint(4 * NXY);
  int32_t *haz = hint.data(), *hau = haz + NXY, *hav = hau + NXY, *hax = hav + NXY;
  memcpy(haz, az, sizeof(int32_t) * NXY);
  std::vector<int32_t> dau(NXY), dav(NXY), dax(NXY);
  masks_from_az(NX, NY, haz, dau.data(), dav.data(), dax.data());
  memcpy(hau, au ? au : dau.data(), sizeof(int32_t) * NXY);
  memcpy(hav, av ? av : dav.data(), sizeof(int32_t) * NXY);
  memcpy(hax, ax ? ax : dax.data(), sizeof(int32_t) * NXY);

  // init_advection, advection.F90:263-318
  std::vector<uint8_t> m(11 * NXY, 0);
  uint8_t *updH = &m[0], *uflux = &m[NXY], *vflux = &m[2 * NXY], *xflux = &m[3 * NXY],
          *uupdU = &m[4 * NXY], *vupdV = &m[5 * NXY], *ufluxU = &m[6 * NXY],
          *vfluxV = &m[7 * NXY];
  for (long n = 0; n < NXY; n++) {
    updH[n] = haz[n] == 1;
    uflux[n] = hau[n] == 1 || hau[n] == 2;
    vflux[n] = hav[n] == 1 || hav[n] == 2;
    xflux[n] = hax[n] == 1;
    uupdU[n] = hau[n] == 1;
    vupdV[n] = hav[n] == 1;
  }
  for (long j = 0; j < NY; j++) uflux[NX - 1 + NX * j] = xflux[NX - 1 + NX * j] = 0;
  for (long i = 0; i < NX; i++) vflux[i + NX * (NY - 1)] = xflux[i + NX * (NY - 1)] = 0;
  for (long j = 0; j < NY; j++)
    for (long i = 0; i < NX - 1; i++) ufluxU[i + NX * j] = updH[i + 1 + NX * j];
  for (long j = 0; j < NY - 1; j++)
    for (long i = 0; i < NX; i++) vfluxV[i + NX * j] = updH[i + NX * (j + 1)];
  // the same masks packed one byte per cell for the strip kernels
  uint8_t *flH = &m[8 * NXY], *flU = &m[9 * NXY], *flV = &m[10 * NXY];
  for (long n = 0; n < NXY; n++) {
    flH[n] = (uint8_t)(uflux[n] * gt::MB_UFLUX | vflux[n] * gt::MB_VFLUX |
                       updH[n] * gt::MB_UUPD | updH[n] * gt::MB_VUPD |
                       (haz[n] == 1) * gt::MB_WCOL);
    flU[n] = (uint8_t)(ufluxU[n] * gt::MB_UFLUX | xflux[n] * gt::MB_VFLUX |
                       uupdU[n] * gt::MB_UUPD | uflux[n] * gt::MB_VUPD |
                       (hau[n] == 1) * gt::MB_WCOL);
    flV[n] = (uint8_t)(xflux[n] * gt::MB_UFLUX | vfluxV[n] * gt::MB_VFLUX |
                       vflux[n] * gt::MB_UUPD | vupdV[n] * gt::MB_VUPD |
                       (hav[n] == 1) * gt::MB_WCOL);
  }

  hipError_t e = hipMalloc(&c->dmask, m.size());
  if (e == hipSuccess) e = hipMalloc(&c->dint, sizeof(int32_t) * 4 * NXY);
  if (e == hipSuccess) e = hipMemcpy(c->dmask, m.data(), m.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(c->dint, hint.data(), sizeof(int32_t) * 4 * NXY, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&c->nwarn, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemset(c->nwarn, 0, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&c->zz, sizeof(double) * c->N3);
  if (e == hipSuccess) e = hipMemset(c->zz, 0, sizeof(double) * c->N3);
  if (e != hipSuccess) {
    gt_destroy(c);
    return set_err(GT_EHIP, "gt_create: %s", hipGetErrorString(e));
  }
  // the device address of the allocation-free entry points' scratch array,
  // looked up once per device with that device current
  if (noalloc_scratch(&c->noalloc) != GT_OK) {
    gt_destroy(c);
    return GT_EHIP;  // the message names the failed lookup
  }
  const uint8_t* dm = c->dmask;
  c->az = c->dint; c->au = c->dint + NXY; c->av = c->dint + 2 * NXY; c->ax = c->dint + 3 * NXY;
  c->H = {dm + NXY, dm + 2 * NXY, dm, dm, dm, c->az, dm + 8 * NXY};
  c->U = {dm + 6 * NXY, dm + 3 * NXY, dm + 4 * NXY, dm + NXY, dm + NXY, c->au, dm + 9 * NXY};
  c->V = {dm + 3 * NXY, dm + 7 * NXY, dm + 2 * NXY, dm + 5 * NXY, dm + 2 * NXY, c->av,
          dm + 10 * NXY};
  *out = c;
  return GT_OK;
}

extern "C" int gt_get_masks(gt_ctx* c, int32_t* az, int32_t* au, int32_t* av, int32_t* ax) {
  if (!c) return set_err(GT_EARG, "gt_get_masks: NULL context");
  int32_t* dst[4] = {az, au, av, ax};
  const size_t bytes = sizeof(int32_t) * (size_t)c->d.NXY;
  for (int q = 0; q < 4; q++)
    if (dst[q])
      HIPCHK(hipMemcpyAsync(dst[q], c->dint + q * c->d.NXY, bytes, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return GT_OK;
}

extern "C" int gt_destroy(gt_ctx* c) {
  if (!c) return GT_OK;
  void* ps[] = {c->dmask, c->dint, c->fi, c->hi, c->adv, c->tmp, c->scr7, c->zz,
                c->P, c->slice3, c->bpart, c->bout, c->nwarn, c->ws, c->dlev, c->gga,
                c->dbad, c->vdw, c->tri, c->mwork, c->w32, c->idp0, c->bbatch};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete c;
  return GT_OK;
}

extern "C" int gt_set_stream(gt_ctx* c, void* s) {
  if (!c) return set_err(GT_EARG, "gt_set_stream: NULL context");
  c->st = (hipStream_t)s;
  return GT_OK;
}

extern "C" int gt_synchronize(gt_ctx* c) {
  if (!c) return set_err(GT_EARG, "gt_synchronize: NULL context");
  HIPCHK(hipStreamSynchronize(c->st));
  return GT_OK;
}

extern "C" int gt_set_adv_ver_iterations(gt_ctx* c, int iters) {
  if (!c) return set_err(GT_EARG, "gt_set_adv_ver_iterations: NULL context");
  c->adv_ver_iterations = iters;
  return GT_OK;
}

extern "C" long long gt_warnings(gt_ctx* c) {
  if (!c) return -1;
  unsigned long long h = 0;
  if (hipMemcpyAsync(&h, c->nwarn, sizeof(h), hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(c->st) != hipSuccess) return -1;
  if (hipMemsetAsync(c->nwarn, 0, sizeof(h), c->st) != hipSuccess) return -1;
  return (long long)h;
}

// ---- helpers ---------------------------------------------------------------
static const GridMasks* grid_for(gt_ctx* c, int tag) {
  switch (tag) {
    case GT_H_TAG: case GT_D_TAG: case GT_W_TAG: return &c->H;
    case GT_U_TAG: return &c->U;
    case GT_V_TAG: return &c->V;
    default: return nullptr;
  }
}
static bool split_scheme(int s) {
  return s == GT_UPSTREAM || s == GT_P2 || s == GT_SUPERBEE || s == GT_MUSCL ||
         s == GT_P2_PDM;
}
static bool twodh_scheme(int s) {
  return s == GT_UPSTREAM_2DH || s == GT_J7 || s == GT_FCT || s == GT_P2_2DH;
}
static int copy_d2d(gt_ctx* c, double* dst, const double* src, long n) {
  if (dst == src || n <= 0) return GT_OK;
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
  return GT_OK;
}
#define TRY(x)                 \
  do {                         \
    int r_ = (x);              \
    if (r_ != GT_OK) return r_; \
  } while (0)
#define LCHK(x)                                                                         \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess) return set_err(GT_EHIP, "launch failed: %s", hipGetErrorString(e_)); \
  } while (0)

// A field advanced by a sequence of sub-steps that read neighbours of the
// current field: every sub-step writes the other buffer, and the result is
// copied back to the caller's array at the end if needed.  `nlev` levels
// starting at level `k0` are sub-stepped.
struct PingPong {
  double* user;
  double* cur;
  double* oth;
  void swap() { double* t = cur; cur = oth; oth = t; }
};

// horizontal sub-steps over `ns` slices starting at level k0
static int step_u(gt_ctx* c, const GridMasks* g, int k0, int ns, const double* f,
                  const double* fis, double* fid, double* Di, double* adv, const double* U,
                  const double* DU, double dt, double sf, int scheme, double AH) {
  long o = c->d.NXY * k0;
  LCHK(gt::launch_split_u(c->st, c->d, ns, f + o, fis + o, fid + o, Di + o, adv + o, U + o,
                          DU + o, g->uflux, g->uupd, dt, sf, scheme, AH, c->dx, c->dy,
                          c->ard1));
  return GT_OK;
}
static int step_v(gt_ctx* c, const GridMasks* g, int k0, int ns, const double* f,
                  const double* fis, double* fid, double* Di, double* adv, const double* V,
                  const double* DV, double dt, double sf, int scheme, double AH) {
  long o = c->d.NXY * k0;
  LCHK(gt::launch_split_v(c->st, c->d, ns, f + o, fis + o, fid + o, Di + o, adv + o, V + o,
                          DV + o, g->vflux, g->vupd, dt, sf, scheme, AH, c->dx, c->dy,
                          c->ard1));
  return GT_OK;
}
static int step_up2dh(gt_ctx* c, const GridMasks* g, int k0, int ns, const double* f,
                      const double* fis, double* fid, double* Di, double* adv, const double* U,
                      const double* V, const double* Dn, double dt, double AH) {
  long o = c->d.NXY * k0;
  LCHK(gt::launch_upstream_2dh(c->st, c->d, ns, f + o, fis + o, fid + o, Di + o, adv + o,
                               U + o, V + o, Dn + o, g->az, dt, AH, c->dx, c->dy, c->ard1));
  return GT_OK;
}
static int step_w(gt_ctx* c, const GridMasks* g, const double* f, const double* fis,
                  double* fid, double* hi, double* adv, const double* ww, double dt, double sf,
                  int scheme, int tag, int itersmax) {
  if (scheme == GT_NOADV) return copy_d2d(c, fid, fis, c->N3);
  int kshift = tag == GT_W_TAG ? 1 : 0;
  if (itersmax > 1) {
    ENSURE(c->scr7, 7 * c->N3);
    LCHK(gt::launch_split_w_iter(c->st, c->d, f, fis, fid, hi, adv, ww, g->az, c->scr7,
                                 c->nwarn, dt, sf, scheme, kshift, itersmax));
  } else {
    LCHK(gt::launch_split_w(c->st, c->d, f, fis, fid, hi, adv, ww, g->az, dt, sf, scheme,
                            kshift));
  }
  return GT_OK;
}

// do_advection (advection.F90:362-702) on ns slices from level k0 of 3-D
// arrays (ns=1, k0=0 for the 2-D entry point).  p_Di/p_adv are already
// initialised by the caller.
static int adv2d_slices(gt_ctx* c, const GridMasks* g, int k0, int ns, double dt, double* f,
                        const double* U, const double* V, const double* DU, const double* DV,
                        const double* Do, const double* Dn, int split, int scheme, double AH,
                        double* pDi, double* padv, double* fi_buf, double* tmp_buf,
                        long nfield) {
  if (scheme == GT_NOADV) return GT_OK;
  const long o = c->d.NXY * k0, n = c->d.NXY * ns;
  if (split == GT_NOSPLIT) {
    if (split_scheme(scheme)) {
      TRY(copy_d2d(c, fi_buf + o, f + o, n));
      TRY(step_u(c, g, k0, ns, f, fi_buf, fi_buf, pDi, padv, U, DU, dt, 1.0, scheme, AH));
      TRY(step_v(c, g, k0, ns, f, fi_buf, fi_buf, pDi, padv, V, DV, dt, 1.0, scheme, AH));
      LCHK(gt::launch_finalise(c->st, c->d, k0, ns, f, Do, padv, pDi, g->fin, dt));
      return GT_OK;
    }
    if (scheme == GT_UPSTREAM_2DH) {
      TRY(step_up2dh(c, g, k0, ns, f, f, tmp_buf, pDi, padv, U, V, Dn, dt, AH));
      return copy_d2d(c, f + o, tmp_buf + o, n);
    }
  } else if (split == GT_FULLSPLIT || split == GT_HALFSPLIT) {
    (void)nfield;
    PingPong pp{f, f, tmp_buf};
    double sfu = split == GT_HALFSPLIT ? 0.5 : 1.0;
    TRY(step_u(c, g, k0, ns, pp.cur, pp.cur, pp.oth, pDi, padv, U, DU, dt, sfu, scheme, AH));
    pp.swap();
    TRY(step_v(c, g, k0, ns, pp.cur, pp.cur, pp.oth, pDi, padv, V, DV, dt, 1.0, scheme, AH));
    pp.swap();
    if (split == GT_HALFSPLIT) {
      TRY(step_u(c, g, k0, ns, pp.cur, pp.cur, pp.oth, pDi, padv, U, DU, dt, 0.5, scheme, AH));
      pp.swap();
    }
    if (pp.cur != f) return copy_d2d(c, f + o, pp.cur + o, n);
    return GT_OK;
  }
  return set_err(GT_EINVAL_SPLIT, "do_advection: split is invalid");
}

static int check_2d(int split, int scheme) {
  if (scheme == GT_NOADV) return GT_OK;
  if (scheme < 0 || scheme > 9) return set_err(GT_EINVAL_SCHEME, "do_advection: scheme is invalid");
  switch (split) {
    case GT_NOSPLIT:
      if (scheme == GT_J7 || scheme == GT_FCT || scheme == GT_P2_2DH)
        return set_err(GT_ENOTIMPL, "do_advection: scheme %d is not part of this build", scheme);
      return GT_OK;
    case GT_FULLSPLIT:
    case GT_HALFSPLIT:
      if (twodh_scheme(scheme))
        return set_err(GT_ESCHEME_SPLIT, "do_advection: scheme not valid for split");
      return GT_OK;
    default:
      return set_err(GT_EINVAL_SPLIT, "do_advection: split is invalid");
  }
}

// descriptor of one domain for the strip kernels; ws holds the work fields
// f2 (ntracers fields of TF), d2 and, with advres, a2
template <typename TF>
static void fill_domain(gt_ctx* c, const GridMasks* g, gt::FusedDomain<TF>* D, const TF* fin,
                        TF* fout, long tstride, int ntracers, const double* uu, const double* vv,
                        const double* ww, const double* hu, const double* hv, const double* ho,
                        double* hires, double* advres) {
  D->d = c->d;
  D->fin = fin; D->fout = fout; D->tstride = tstride;
  char* w = (char*)c->ws;
  D->d2 = (double*)w;
  D->a2 = advres ? (double*)w + c->N3 : nullptr;
  D->f2 = (TF*)((double*)w + (advres ? 2 : 1) * c->N3);
  (void)ntracers;
  D->uu = uu; D->vv = vv; D->ww = ww; D->hu = hu; D->hv = hv; D->ho = ho;
  D->flags = g->flags;
  D->hires = hires; D->advres = advres;
  D->dx = c->dx; D->dy = c->dy; D->ard1 = c->ard1;
  D->bpart = nullptr;
}

// the strip kernels address one field of one tracer with 32-bit byte offsets
static bool strip_fits(const gt_ctx* c) {
  return (unsigned long long)c->N3 * sizeof(double) < (1ull << 32);
}

// bytes of work fields a domain needs for fill_domain
static size_t strip_ws_bytes(const gt_ctx* c, size_t tf_bytes, int ntracers, bool adv) {
  return sizeof(double) * (size_t)c->N3 * (adv ? 2 : 1) + tf_bytes * (size_t)c->N3 * ntracers;
}

// strip-levels of the second strip kernel at its smallest strip height (8
// rows): the most budget partials a tuning can ask for
static long budget_strip_levels(const gt_ctx* c) {
  return (long)((c->d.NX + gt::kStripOut - 1) / gt::kStripOut) * ((c->d.NY + 7) / 8) * c->kmax;
}
static size_t budget_ws_bytes(const gt_ctx* c, int ntracers) {
  return sizeof(double) * 2 * (size_t)(ntracers + 1) * (size_t)budget_strip_levels(c);
}

// ---- advection entry points --------------------------------------------------
__global__ void k_reconstruct_batch(int scheme, long n, const double* cfl, const double* fuu,
                                    const double* fu, const double* fd, double* out) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) out[q] = gt::reconstruct<double>(scheme, cfl[q], fuu[q], fu[q], fd[q]);
}

extern "C" int gt_adv_interfacial_reconstruction(gt_ctx* c, int scheme, long n,
                                                 const double* cfl, const double* fuu,
                                                 const double* fu, const double* fd,
                                                 double* out) {
  if (!c || n < 0) return set_err(GT_EARG, "gt_adv_interfacial_reconstruction: bad argument");
  if (n == 0) return GT_OK;
  hipLaunchKernelGGL(k_reconstruct_batch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->st,
                     scheme, n, cfl, fuu, fu, fd, out);
  LCHK(hipGetLastError());
  return GT_OK;
}

static int split_u_impl(gt_ctx* c, const GridMasks* g, double dt, const double* f, double* fi,
                        double* Di, double* adv, const double* U, const double* DU,
                        double splitfac, int scheme, double AH) {
  if (!split_scheme(scheme)) return set_err(GT_EINVAL_SCHEME, "adv_split_u: scheme is invalid");
  if (f == fi) {
    ENSURE(c->slice3, 3 * c->d.NXY);
    TRY(step_u(c, g, 0, 1, f, f, c->slice3, Di, adv, U, DU, dt, splitfac, scheme, AH));
    return copy_d2d(c, fi, c->slice3, c->d.NXY);
  }
  return step_u(c, g, 0, 1, f, fi, fi, Di, adv, U, DU, dt, splitfac, scheme, AH);
}

extern "C" int gt_adv_split_u(gt_ctx* c, double dt, const double* f, double* fi, double* Di,
                              double* adv, const double* U, const double* DU, double splitfac,
                              int scheme, double AH, int tag) {
  if (!c) return set_err(GT_EARG, "gt_adv_split_u: NULL context");
  const GridMasks* g = grid_for(c, tag);
  if (!g) return set_err(GT_EINVAL_TAG, "adv_split_u: tag is invalid");
  return split_u_impl(c, g, dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH);
}

static int split_v_impl(gt_ctx* c, const GridMasks* g, double dt, const double* f, double* fi,
                        double* Di, double* adv, const double* V, const double* DV,
                        double splitfac, int scheme, double AH) {
  if (!split_scheme(scheme)) return set_err(GT_EINVAL_SCHEME, "adv_split_v: scheme is invalid");
  if (f == fi) {
    ENSURE(c->slice3, 3 * c->d.NXY);
    TRY(step_v(c, g, 0, 1, f, f, c->slice3, Di, adv, V, DV, dt, splitfac, scheme, AH));
    return copy_d2d(c, fi, c->slice3, c->d.NXY);
  }
  return step_v(c, g, 0, 1, f, fi, fi, Di, adv, V, DV, dt, splitfac, scheme, AH);
}

extern "C" int gt_adv_split_v(gt_ctx* c, double dt, const double* f, double* fi, double* Di,
                              double* adv, const double* V, const double* DV, double splitfac,
                              int scheme, double AH, int tag) {
  if (!c) return set_err(GT_EARG, "gt_adv_split_v: NULL context");
  const GridMasks* g = grid_for(c, tag);
  if (!g) return set_err(GT_EINVAL_TAG, "adv_split_v: tag is invalid");
  return split_v_impl(c, g, dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH);
}

static int split_w_impl(gt_ctx* c, const GridMasks* g, double dt, const double* f, double* fi,
                        double* hi, double* adv, const double* ww, double splitfac, int scheme,
                        int tag, int itersmax) {
  if (scheme != GT_NOADV && !split_scheme(scheme))
    return set_err(GT_EINVAL_SCHEME, "adv_split_w: scheme is invalid");
  if (scheme == GT_NOADV) return GT_OK;
  if (f == fi) {
    ENSURE(c->tmp, c->N3);
    TRY(step_w(c, g, f, f, c->tmp, hi, adv, ww, dt, splitfac, scheme, tag, itersmax));
    return copy_d2d(c, fi, c->tmp, c->N3);
  }
  return step_w(c, g, f, fi, fi, hi, adv, ww, dt, splitfac, scheme, tag, itersmax);
}

extern "C" int gt_adv_split_w(gt_ctx* c, double dt, const double* f, double* fi, double* hi,
                              double* adv, const double* ww, double splitfac, int scheme,
                              int tag, int itersmax) {
  if (!c) return set_err(GT_EARG, "gt_adv_split_w: NULL context");
  // adv_split_w receives adv_grid%az; for the 3-D tags that is the H grid
  const GridMasks* g = grid_for(c, tag);
  if (!g) return set_err(GT_EINVAL_TAG, "adv_split_w: tag is invalid");
  return split_w_impl(c, g, dt, f, fi, hi, adv, ww, splitfac, scheme, tag, itersmax);
}

static int upstream_2dh_impl(gt_ctx* c, const GridMasks* g, double dt, const double* f,
                             double* fi, double* Di, double* adv, const double* U,
                             const double* V, const double* Dn, double AH) {
  if (f == fi) {
    ENSURE(c->slice3, 3 * c->d.NXY);
    TRY(step_up2dh(c, g, 0, 1, f, f, c->slice3, Di, adv, U, V, Dn, dt, AH));
    return copy_d2d(c, fi, c->slice3, c->d.NXY);
  }
  return step_up2dh(c, g, 0, 1, f, fi, fi, Di, adv, U, V, Dn, dt, AH);
}

extern "C" int gt_adv_upstream_2dh(gt_ctx* c, double dt, const double* f, double* fi,
                                   double* Di, double* adv, const double* U, const double* V,
                                   const double* Dn, const double* DU, const double* DV,
                                   double AH, int tag) {
  (void)DU; (void)DV;
  if (!c) return set_err(GT_EARG, "gt_adv_upstream_2dh: NULL context");
  const GridMasks* g = grid_for(c, tag);
  if (!g) return set_err(GT_EINVAL_TAG, "adv_upstream_2dh: tag is invalid");
  return upstream_2dh_impl(c, g, dt, f, fi, Di, adv, U, V, Dn, AH);
}

// The same four routines with the masks the reference's explicit interfaces
// take as arguments (advection.F90:91-194, advection_3d.F90:55-68): the
// caller's t_adv_grid masks as bytes (1 = .true.) over E2DFIELD, and az
// (adv_grid%az) as int32, all device memory.
extern "C" int gt_adv_split_u_m(gt_ctx* c, double dt, const double* f, double* fi, double* Di,
                                double* adv, const double* U, const double* DU,
                                double splitfac, int scheme, double AH,
                                const uint8_t* mask_flux, const uint8_t* mask_update) {
  if (!c || !mask_flux || !mask_update) return set_err(GT_EARG, "gt_adv_split_u_m: NULL argument");
  GridMasks g{};
  g.uflux = mask_flux;
  g.uupd = mask_update;
  return split_u_impl(c, &g, dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH);
}

extern "C" int gt_adv_split_v_m(gt_ctx* c, double dt, const double* f, double* fi, double* Di,
                                double* adv, const double* V, const double* DV,
                                double splitfac, int scheme, double AH,
                                const uint8_t* mask_flux, const uint8_t* mask_update) {
  if (!c || !mask_flux || !mask_update) return set_err(GT_EARG, "gt_adv_split_v_m: NULL argument");
  GridMasks g{};
  g.vflux = mask_flux;
  g.vupd = mask_update;
  return split_v_impl(c, &g, dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH);
}

extern "C" int gt_adv_split_w_m(gt_ctx* c, double dt, const double* f, double* fi, double* hi,
                                double* adv, const double* ww, double splitfac, int scheme,
                                int tag, const int32_t* az, int itersmax) {
  if (!c || !az) return set_err(GT_EARG, "gt_adv_split_w_m: NULL argument");
  if (!grid_for(c, tag)) return set_err(GT_EINVAL_TAG, "adv_split_w: tag is invalid");
  GridMasks g{};
  g.az = az;
  return split_w_impl(c, &g, dt, f, fi, hi, adv, ww, splitfac, scheme, tag, itersmax);
}

extern "C" int gt_adv_upstream_2dh_m(gt_ctx* c, double dt, const double* f, double* fi,
                                     double* Di, double* adv, const double* U, const double* V,
                                     const double* Dn, const double* DU, const double* DV,
                                     double AH, const int32_t* az) {
  (void)DU; (void)DV;
  if (!c || !az) return set_err(GT_EARG, "gt_adv_upstream_2dh_m: NULL argument");
  GridMasks g{};
  g.az = az;
  return upstream_2dh_impl(c, &g, dt, f, fi, Di, adv, U, V, Dn, AH);
}

extern "C" int gt_do_advection(gt_ctx* c, double dt, double* f, const double* U,
                               const double* V, const double* DU, const double* DV,
                               const double* Do, const double* Dn, int split, int scheme,
                               double AH, int tag, double* Dires, double* advres) {
  if (!c) return set_err(GT_EARG, "gt_do_advection: NULL context");
  if (tag == GT_W_TAG) return set_err(GT_EINVAL_TAG, "do_advection: tag is invalid");
  const GridMasks* g = grid_for(c, tag);
  if (!g) return set_err(GT_EINVAL_TAG, "do_advection: tag is invalid");
  TRY(check_2d(split, scheme));
  const long n = c->d.NXY;
  ENSURE(c->slice3, 3 * n);
  double* fi = c->slice3;
  double* tmp = c->slice3 + n;
  double* pDi = Dires;
  double* padv = advres;
  if (!pDi || !padv) {
    ENSURE(c->hi, c->N3);
    ENSURE(c->adv, c->N3);
    if (!pDi) pDi = c->hi;
    if (!padv) padv = c->adv;
  }
  TRY(copy_d2d(c, pDi, Do, n));
  HIPCHK(hipMemsetAsync(padv, 0, sizeof(double) * n, c->st));
  return adv2d_slices(c, g, 0, 1, dt, f, U, V, DU, DV, Do, Dn, split, scheme, AH, pDi, padv, fi,
                      tmp, n);
}

extern "C" int gt_do_advection_3d(gt_ctx* c, double dt, double* f, const double* uu,
                                  const double* vv, const double* ww, const double* hu,
                                  const double* hv, const double* ho, const double* hn,
                                  int split, int hscheme, int vscheme, double AH, int tag,
                                  double* hires, double* advres) {
  if (!c) return set_err(GT_EARG, "gt_do_advection_3d: NULL context");
  int tag2d;
  const GridMasks* g;
  switch (tag) {
    case GT_H_TAG: case GT_D_TAG: tag2d = GT_H_TAG; g = &c->H; break;
    case GT_U_TAG: tag2d = GT_U_TAG; g = &c->U; break;
    case GT_V_TAG: tag2d = GT_V_TAG; g = &c->V; break;
    case GT_W_TAG: tag2d = GT_H_TAG; g = &c->H; break;
    default: return set_err(GT_EINVAL_TAG, "do_advection_3d: tag is invalid");
  }
  (void)tag2d;
  const bool hadv = hscheme != GT_NOADV, vadv = vscheme != GT_NOADV;
  // the checks print_adv_settings_3d makes at start-up (advection_3d.F90:
  // 600-652) plus the stops of do_advection_3d, made before any work
  if (hscheme < 0 || hscheme > 9)
    return set_err(GT_EINVAL_SCHEME, "do_advection_3d: hscheme is invalid");
  if (vadv && !split_scheme(vscheme))
    return set_err(GT_EINVAL_SCHEME, "do_advection_3d: vscheme %d is invalid", vscheme);
  if (hadv && vadv) {
    if (split < 0 || split > 3) return set_err(GT_EINVAL_SPLIT, "do_advection_3d: split is invalid");
    if ((split == GT_FULLSPLIT || split == GT_HALFSPLIT || split == GT_HVSPLIT) &&
        twodh_scheme(hscheme) && !(split == GT_HVSPLIT && hscheme == GT_UPSTREAM_2DH))
      return set_err(split == GT_HVSPLIT ? GT_ENOTIMPL : GT_ESCHEME_SPLIT,
                     "do_advection_3d: hscheme not valid for split");
    if (split == GT_NOSPLIT && (hscheme == GT_J7 || hscheme == GT_FCT || hscheme == GT_P2_2DH))
      return set_err(GT_ENOTIMPL, "do_advection_3d: hscheme %d is not part of this build",
                     hscheme);
  } else if (hadv) {
    TRY(check_2d(split, hscheme));
  }
  const long N3 = c->N3, NXY = c->d.NXY;
  const int kmax = c->kmax;
  double* phi = hires;
  double* padv = advres;
  if (!phi) { ENSURE(c->hi, N3); phi = c->hi; }
  if (!padv) { ENSURE(c->adv, N3); padv = c->adv; }
  TRY(copy_d2d(c, phi, ho, N3));
  HIPCHK(hipMemsetAsync(padv, 0, sizeof(double) * N3, c->st));
  if (!hadv && !vadv) return GT_OK;
  ENSURE(c->tmp, N3);
  ENSURE(c->fi, N3);
  double* tmp = c->tmp;
  // level 0 is never advanced horizontally; keep it equal in both buffers
  TRY(copy_d2d(c, tmp, f, NXY));
  const int itv = c->adv_ver_iterations;

  if (!hadv) {
    if (kmax > 1) {
      TRY(step_w(c, g, f, f, tmp, phi, padv, ww, dt, 1.0, vscheme, tag, itv));
      TRY(copy_d2d(c, f, tmp, N3));
    }
    return GT_OK;
  }
  if (!vadv) {
    return adv2d_slices(c, g, 1, kmax, dt, f, uu, vv, hu, hv, ho, hn, split, hscheme, AH, phi,
                        padv, c->fi, tmp, N3);
  }
  switch (split) {
    case GT_NOSPLIT: {
      double* fi = c->fi;
      TRY(copy_d2d(c, fi, f, N3));
      if (split_scheme(hscheme)) {
        TRY(step_u(c, g, 1, kmax, f, fi, fi, phi, padv, uu, hu, dt, 1.0, hscheme, AH));
        TRY(step_v(c, g, 1, kmax, f, fi, fi, phi, padv, vv, hv, dt, 1.0, hscheme, AH));
      } else {  // UPSTREAM_2DH
        TRY(step_up2dh(c, g, 1, kmax, f, fi, fi, phi, padv, uu, vv, hn, dt, AH));
      }
      if (kmax > 1) TRY(step_w(c, g, f, fi, fi, phi, padv, ww, dt, 1.0, vscheme, tag, 1));
      LCHK(gt::launch_finalise(c->st, c->d, 1, kmax, f, ho, padv, phi, g->fin, dt));
      return GT_OK;
    }
    case GT_FULLSPLIT: {
      PingPong pp{f, f, tmp};
      TRY(step_u(c, g, 1, kmax, pp.cur, pp.cur, pp.oth, phi, padv, uu, hu, dt, 1.0, hscheme, AH));
      pp.swap();
      TRY(step_v(c, g, 1, kmax, pp.cur, pp.cur, pp.oth, phi, padv, vv, hv, dt, 1.0, hscheme, AH));
      pp.swap();
      if (kmax > 1) {
        TRY(step_w(c, g, pp.cur, pp.cur, pp.oth, phi, padv, ww, dt, 1.0, vscheme, tag, itv));
        pp.swap();
      }
      if (pp.cur != f) TRY(copy_d2d(c, f, pp.cur, N3));
      return GT_OK;
    }
    case GT_HALFSPLIT: {
      if (c->fused && itv <= 1 && kmax > 1 && gt::fused_supported(hscheme, vscheme) &&
          strip_fits(c)) {
        // f is read at level 0 only by the second kernel, which writes the
        // same values there, so the result goes to f directly
        TRY(ensure_ws(c, strip_ws_bytes(c, sizeof(double), 1, advres != nullptr)));
        gt::FusedArgs<double> A{};
        fill_domain<double>(c, g, &A.doms[0], f, f, 0, 1, uu, vv, ww, hu, hv, ho, hires, advres);
        A.ndom = 1;
        A.ntracers = 1;
        gt::strip_rows(c->tile, A.rows);
        A.dt = dt; A.AH = AH;
        A.hscheme = hscheme; A.vscheme = vscheme; A.kshift = tag == GT_W_TAG ? 1 : 0;
        gt::strip_plan(A, 1);
        LCHK(gt::launch_fused_f64(c->st, A, 1, advres != nullptr));
        return GT_OK;
      }
      PingPong pp{f, f, tmp};
      TRY(step_u(c, g, 1, kmax, pp.cur, pp.cur, pp.oth, phi, padv, uu, hu, dt, 0.5, hscheme, AH));
      pp.swap();
      TRY(step_v(c, g, 1, kmax, pp.cur, pp.cur, pp.oth, phi, padv, vv, hv, dt, 0.5, hscheme, AH));
      pp.swap();
      if (kmax > 1) {
        TRY(step_w(c, g, pp.cur, pp.cur, pp.oth, phi, padv, ww, dt, 1.0, vscheme, tag, itv));
        pp.swap();
      }
      TRY(step_v(c, g, 1, kmax, pp.cur, pp.cur, pp.oth, phi, padv, vv, hv, dt, 0.5, hscheme, AH));
      pp.swap();
      TRY(step_u(c, g, 1, kmax, pp.cur, pp.cur, pp.oth, phi, padv, uu, hu, dt, 0.5, hscheme, AH));
      pp.swap();
      if (pp.cur != f) TRY(copy_d2d(c, f, pp.cur, N3));
      return GT_OK;
    }
    case GT_HVSPLIT: {
      TRY(adv2d_slices(c, g, 1, kmax, dt, f, uu, vv, hu, hv, ho, hn, GT_NOSPLIT, hscheme, AH, phi,
                       padv, c->fi, tmp, N3));
      if (kmax > 1) {
        TRY(step_w(c, g, f, f, tmp, phi, padv, ww, dt, 1.0, vscheme, tag, itv));
        TRY(copy_d2d(c, f, tmp, N3));
      }
      return GT_OK;
    }
  }
  return set_err(GT_EINVAL_SPLIT, "do_advection_3d: split is invalid");
}

// ---- input producers (m3d.F90:446-548) ---------------------------------------
extern "C" int gt_start_macro(gt_ctx* c, int M, double min_depth, const double* H,
                              const double* HU, const double* HV, const double* z, double* sseo,
                              double* ssen, double* ssuo, double* ssun, double* ssvo,
                              double* ssvn, double* Dn, double* Dun, double* Dvn, double* Uint,
                              double* Vint) {
  const void* ps[] = {H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn, Dn, Dun, Dvn, Uint, Vint};
  if (!c) return set_err(GT_EARG, "gt_start_macro: NULL context");
  for (const void* p : ps)
    if (!p) return set_err(GT_EARG, "gt_start_macro: NULL field");
  if (M < 1) return set_err(GT_EARG, "gt_start_macro: M must be >= 1");
  const double split = 1.0 / (double)(float)M;  // _ONE_/float(M), start_macro.F90:84
  LCHK(gt::launch_start_macro(c->st, c->d, split, min_depth, H, HU, HV, z, sseo, ssen, ssuo,
                              ssun, ssvo, ssvn, Dn, Dun, Dvn, Uint, Vint));
  return GT_OK;
}

extern "C" int gt_set_coordinates(gt_ctx* c, int vert_cord, double ddu, double ddl,
                                  double d_gamma, int gamma_surf, double min_depth) {
  if (!c) return set_err(GT_EARG, "gt_set_coordinates: NULL context");
  if (vert_cord == 4)
    return set_err(GT_ENOTIMPL, "coordinates(): hybrid_coordinates not coded yet");
  if (vert_cord < 1 || vert_cord > 3)
    return set_err(GT_ENOTIMPL, "gt_set_coordinates: vert_cord %d is not in this build",
                   vert_cord);
  c->vert_cord = vert_cord;
  c->ddu = ddu;
  c->ddl = ddl;
  c->d_gamma = d_gamma;
  c->gamma_surf = gamma_surf != 0;
  c->min_depth = min_depth;
  c->coord_first = 1;
  c->equiv_sigma = 0;
  return GT_OK;
}

// the surface/bottom zoom, sigma_coordinates.F90:72-76, general_coordinates.F90:72-75
static double zoom_level(double ddu, double ddl, int k, int kmax) {
  double g = tanh((ddl + ddu) * k / (double)(float)kmax - ddl) + tanh(ddl);
  return g / (tanh(ddl) + tanh(ddu)) - 1.0;
}

// the first call of sigma_coordinates / general_coordinates: the level
// arrays on the host, as the reference computes them, then on the device
static int coord_levels(gt_ctx* c) {
  const int kmax = c->kmax, K1 = kmax + 1;
  std::vector<double> lev(2 * (size_t)K1, 0.0);
  if (c->vert_cord == 1) {
    if (c->ddu <= 0.0 && c->ddl <= 0.0) {
      c->equiv_sigma = 1;
      return GT_OK;
    }
    if (c->ddu < 0.0) c->ddu = 0.0;
    if (c->ddl < 0.0) c->ddl = 0.0;
    double gprev = -1.0;  // ga(0)
    lev[0] = 0.0;         // dga(0)
    for (int k = 1; k <= kmax; k++) {
      const double g = zoom_level(c->ddu, c->ddl, k, kmax);
      lev[k] = g - gprev;
      gprev = g;
    }
  } else {
    double* sig = lev.data();
    double* be = lev.data() + K1;
    be[0] = -1.0;
    sig[0] = -1.0;
    if (c->ddu < 0.0) c->ddu = 0.0;
    if (c->ddl < 0.0) c->ddl = 0.0;
    for (int k = 1; k <= kmax; k++) {
      be[k] = zoom_level(c->ddu, c->ddl, k, kmax);
      sig[k] = (double)((float)k / (float)kmax) - 1.0;  // k/float(kmax): an FP32 quotient
    }
  }
  ENSURE(c->dlev, 2 * (long)K1);
  HIPCHK(hipMemcpy(c->dlev, lev.data(), sizeof(double) * lev.size(), hipMemcpyHostToDevice));
  return GT_OK;
}

extern "C" int gt_coordinates(gt_ctx* c, double dt, double cord_relax, double maxdepth,
                              const double* H, const double* HU, const double* HV,
                              const double* sseo, const double* ssen, const double* ssuo,
                              const double* ssun, const double* ssvo, const double* ssvn,
                              double* ho, double* hn, double* huo, double* hun, double* hvo,
                              double* hvn, double* zc) {
  if (!c) return set_err(GT_EARG, "gt_coordinates: NULL context");
  if (!c->vert_cord) return set_err(GT_EARG, "gt_coordinates: call gt_set_coordinates first");
  if (!H || !hn) return set_err(GT_EARG, "gt_coordinates: NULL field");
  if (c->vert_cord == 2) {  // z-levels: nothing but zc, coordinates.F90:71,90,103-107
    c->coord_first = 0;
    if (zc) LCHK(gt::launch_zc(c->st, c->d, H, hn, zc));
    return GT_OK;
  }
  const void* ps[] = {HU, HV, sseo, ssen, ssuo, ssun, ssvo, ssvn, ho, huo, hun, hvo, hvn};
  for (const void* p : ps)
    if (!p) return set_err(GT_EARG, "gt_coordinates: NULL field");
  if (c->vert_cord == 3 && (dt == 0.0 || maxdepth == 0.0))
    return set_err(GT_EARG, "gt_coordinates: dt and maxdepth must be nonzero");
  if (c->coord_first) TRY(coord_levels(c));
  if (c->vert_cord == 1) {
    c->coord_first = 0;
    LCHK(gt::launch_sigma(c->st, c->d, c->equiv_sigma != 0, 1.0 / (double)(float)c->kmax,
                          c->dlev, H, HU, HV, sseo, ssen, ssuo, ssun, ssvo, ssvn, ho, hn, huo,
                          hun, hvo, hvn, zc));
    return GT_OK;
  }
  if (c->coord_first) {
    ENSURE(c->gga, c->N3);
    if (!c->dbad) HIPCHK(hipMalloc(&c->dbad, sizeof(int)));
    HIPCHK(hipMemsetAsync(c->dbad, 0, sizeof(int), c->st));
    const int kk = c->gamma_surf ? c->kmax : 1;
    LCHK(gt::launch_gvc_init(c->st, c->d, kk, c->d_gamma, c->min_depth, c->dlev, H, HU, HV, sseo,
                             ssuo, ssvo, c->gga, hn, huo, hun, hvo, hvn, c->dbad));
    int bad = 0;
    HIPCHK(hipMemcpyAsync(&bad, c->dbad, sizeof(int), hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    if (bad)
      return set_err(GT_EARG, "coordinates: general vertical coordinates not monotone "
                              "(general_coordinates.F90:92-98)");
    c->coord_first = 0;
  }
  LCHK(gt::launch_gvc_step(c->st, c->d, cord_relax / dt, maxdepth, c->d_gamma, c->gga, H, HU, HV,
                           ssen, ssun, ssvn, ho, hn, huo, hun, hvo, hvn, zc));
  return GT_OK;
}

extern "C" int gt_ww_momentum_3d(gt_ctx* c, double dt, const double* uu, const double* vv,
                                 const double* ho, const double* hn, double* ww) {
  if (!c || !uu || !vv || !ho || !hn || !ww)
    return set_err(GT_EARG, "gt_ww_momentum_3d: NULL argument");
  LCHK(gt::launch_ww_momentum(c->st, c->d, dt, c->dx, c->dy, c->ard1, c->az, uu, vv, ho, hn,
                              ww));
  return GT_OK;
}

// ---- restart fields (gt_restart.cpp does the file format) ---------------------
static int restart_host_copies(gt_ctx* c, int n, const gt_restart_field* f,
                               std::vector<std::vector<double>>& host,
                               std::vector<gt_restart_field>& hf, bool download) {
  if (!c || n < 0 || (n > 0 && !f)) return set_err(GT_EARG, "restart: bad argument");
  host.resize((size_t)n);
  hf.resize((size_t)n);
  const long N2 = c->d.NXY;
  HIPCHK(hipStreamSynchronize(c->st));
  for (int v = 0; v < n; v++) {
    if (!f[v].data || (f[v].ndim != 2 && f[v].ndim != 3))
      return set_err(GT_EARG, "restart: field %s: bad data or ndim", f[v].name ? f[v].name : "?");
    const long cnt = f[v].ndim == 3 ? c->N3 : N2;
    host[v].assign((size_t)cnt, 0.0);
    if (download)
      HIPCHK(hipMemcpy(host[v].data(), f[v].data, sizeof(double) * (size_t)cnt,
                       hipMemcpyDeviceToHost));
    hf[v] = f[v];
    hf[v].data = host[v].data();
  }
  return GT_OK;
}

extern "C" int gt_write_restart(gt_ctx* c, const char* path, int with_halos,
                                const gt_restart_time* t, int n, const gt_restart_field* f) {
  std::vector<std::vector<double>> host;
  std::vector<gt_restart_field> hf;
  TRY(restart_host_copies(c, n, f, host, hf, true));
  return gt_restart_write_host(path, c->imin, c->imax, c->jmin, c->jmax, c->kmax, with_halos,
                               t, n, hf.data());
}

extern "C" int gt_read_restart(gt_ctx* c, const char* path, int with_halos, gt_restart_time* t,
                               int n, const gt_restart_field* f) {
  std::vector<std::vector<double>> host;
  std::vector<gt_restart_field> hf;
  // start from the device arrays, so the cells the file does not hold stay
  TRY(restart_host_copies(c, n, f, host, hf, true));
  TRY(gt_restart_read_host(path, c->imin, c->imax, c->jmin, c->jmax, c->kmax, with_halos, t,
                           n, hf.data()));
  for (int v = 0; v < n; v++)
    HIPCHK(hipMemcpy(f[v].data, host[v].data(), sizeof(double) * host[v].size(),
                     hipMemcpyHostToDevice));
  return GT_OK;
}

// ---- a caller of the path: momentum advection --------------------------------
extern "C" int gt_uv_advect_3d(gt_ctx* c, double dt, const double* uu, const double* vv,
                               const double* ww, const double* ho, const double* hn,
                               const double* hun, const double* hvn, int split, int hscheme,
                               int vscheme, double* uuEx, double* vvEx) {
  const void* ps[] = {uu, vv, ww, ho, hn, hun, hvn, uuEx, vvEx};
  if (!c) return set_err(GT_EARG, "gt_uv_advect_3d: NULL context");
  for (const void* p : ps)
    if (!p) return set_err(GT_EARG, "gt_uv_advect_3d: NULL argument");
  if (hscheme == GT_J7) return set_err(GT_ENOTIMPL, "uv_advect_3d: J7 is not in this build");
  if (hscheme == GT_NOADV && vscheme == GT_NOADV) return GT_OK;
  ENSURE(c->mwork, 6 * c->N3);
  double* w[6];
  for (int q = 0; q < 6; q++) w[q] = c->mwork + q * c->N3;
  // U control volumes (uv_advect_3d.F90:62-176), then V (:236-300)
  LCHK(gt::launch_uv_prep(c->st, c->d, true, uu, vv, ww, ho, hn, hun, hvn, c->mwork));
  TRY(gt_do_advection_3d(c, dt, w[0], w[1], w[2], w[3], w[4], w[5], hun, hun, split, hscheme,
                         vscheme, 0.0, GT_U_TAG, nullptr, uuEx));
  LCHK(gt::launch_uv_prep(c->st, c->d, false, uu, vv, ww, ho, hn, hun, hvn, c->mwork));
  TRY(gt_do_advection_3d(c, dt, w[0], w[1], w[2], w[3], w[4], w[5], hvn, hvn, split, hscheme,
                         vscheme, 0.0, GT_V_TAG, nullptr, vvEx));
  return GT_OK;
}

// ---- the step after advection: vertical diffusion --------------------------
extern "C" int gt_salinity_diffusion(gt_ctx* c, double dt, double cnpar, const double* hn,
                                     const double* nuh, double* S) {
  if (!c || !hn || !nuh || !S) return set_err(GT_EARG, "gt_salinity_diffusion: NULL argument");
  if (c->kmax <= 1) return GT_OK;
  ENSURE(c->vdw, 3 * c->N3);
  LCHK(gt::launch_vertical_diffusion(c->st, c->d, false, dt, cnpar, c->az, hn, nuh, S, c->vdw,
                                     nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                     nullptr, -1.0, 0.0, nullptr));
  return GT_OK;
}

extern "C" int gt_temperature_diffusion(gt_ctx* c, double dt, double cnpar, const double* hn,
                                        const double* nuh, const double* swr, const double* shf,
                                        const double* A, const double* g1, const double* g2,
                                        const double* dry_z, const double* S,
                                        double swr_bot_refl_frac, double swr_min_bot_frac,
                                        double* rad, double* T) {
  const void* ps[] = {hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T};
  if (!c) return set_err(GT_EARG, "gt_temperature_diffusion: NULL context");
  for (const void* p : ps)
    if (!p) return set_err(GT_EARG, "gt_temperature_diffusion: NULL argument");
  ENSURE(c->vdw, 3 * c->N3);
  LCHK(gt::launch_vertical_diffusion(c->st, c->d, true, dt, cnpar, c->az, hn, nuh, T, c->vdw,
                                     swr, shf, A, g1, g2, dry_z, S, swr_bot_refl_frac,
                                     swr_min_bot_frac, rad));
  return GT_OK;
}

extern "C" int gt_tridiagonal_batch(gt_ctx* c, long n, int fi, int lt, const double* au,
                                    const double* bu, const double* cu, const double* du,
                                    double* value) {
  if (!c || n < 0 || fi < 0 || lt <= fi)
    return set_err(GT_EARG, "gt_tridiagonal_batch: need n >= 0 and 0 <= fi < lt");
  if (n == 0) return GT_OK;
  if (!au || !bu || !cu || !du || !value)
    return set_err(GT_EARG, "gt_tridiagonal_batch: NULL argument");
  const long need = 2 * n * (long)(lt + 1);
  if (c->tri_n < need) {
    if (c->tri) {
      HIPCHK(hipDeviceSynchronize());  // the old work may still be in use
      HIPCHK(hipFree(c->tri));
      c->tri = nullptr;
      c->tri_n = 0;
    }
    HIPCHK(hipMalloc(&c->tri, sizeof(double) * (size_t)need));
    c->tri_n = need;
  }
  LCHK(gt::launch_tridiagonal(c->st, n, fi, lt, au, bu, cu, du, value, c->tri));
  return GT_OK;
}

// ---- internal pressure -------------------------------------------------------
extern "C" int gt_eos_linear(gt_ctx* c, const double* T, const double* S, double* rho,
                             double* buoy, double T0, double S0, double dtr0, double dsr0) {
  if (!c || !T || !S || !buoy) return set_err(GT_EARG, "gt_eos_linear: NULL argument");
  LCHK(gt::launch_eos_linear(c->st, c->d, T, S, rho, buoy, T0, S0, dtr0, dsr0));
  return GT_OK;
}

extern "C" int gt_do_eqstate(gt_ctx* c, int method, int nonnegsalt, const double* T,
                             const double* S, const double* hn, double* rho, double* buoy,
                             double* alpha, double* beta, double T0, double S0, double dtr0,
                             double dsr0) {
  if (!c || !T || !S || !rho || !buoy) return set_err(GT_EARG, "gt_do_eqstate: NULL argument");
  if (method < 1 || method > 3)
    return set_err(GT_EARG, "init_eqstate(): not a valid eqstate_method");
  if (!alpha != !beta) return set_err(GT_EARG, "gt_do_eqstate: give both alpha and beta or neither");
  if (method != 1 && alpha && !hn) return set_err(GT_EARG, "gt_do_eqstate: hn is needed for alpha/beta");
  LCHK(gt::launch_eqstate(c->st, c->d, method, nonnegsalt != 0, c->az, T, S, hn, rho, buoy,
                          alpha, beta, T0, S0, dtr0, dsr0));
  return GT_OK;
}

static int eos_points(gt_ctx* c, int fn, long n, const double* s, const double* th,
                      const double* p, double* o1, double* o2, int nonneg, const char* who) {
  if (!c || n < 0 || (n > 0 && (!s || !th || !o1 || (fn == 2 && !o2))))
    return set_err(GT_EARG, "%s: bad argument", who);
  LCHK(gt::launch_eos_points(c->st, fn, n, s, th, p, o1, o2, nonneg != 0));
  return GT_OK;
}

extern "C" int gt_rho_from_theta_unesco80(gt_ctx* c, long n, const double* T, const double* S,
                                          double* rho, int nonnegsalt) {
  return eos_points(c, 0, n, S, T, nullptr, rho, nullptr, nonnegsalt,
                    "gt_rho_from_theta_unesco80");
}

extern "C" int gt_rho_from_theta(gt_ctx* c, long n, const double* s, const double* th,
                                 const double* p, double* dens0, double* densp) {
  return eos_points(c, 1, n, s, th, p, dens0, densp, 0, "gt_rho_from_theta");
}

extern "C" int gt_eosall_from_theta(gt_ctx* c, long n, const double* s, const double* th,
                                    const double* p, double* rho_s, double* rho_th) {
  return eos_points(c, 2, n, s, th, p, rho_s, rho_th, 0, "gt_eosall_from_theta");
}

extern "C" int gt_ip_blumberg_mellor(gt_ctx* c, const double* H, const double* hn,
                                     const double* hun, const double* hvn, const double* buoy,
                                     double* zz, double* idpdx, double* idpdy) {
  if (!c) return set_err(GT_EARG, "gt_ip_blumberg_mellor: NULL context");
  if (c->kmax < 2) return set_err(GT_EARG, "ip_blumberg_mellor: kmax must be >= 2");
  LCHK(gt::launch_ip_bm(c->st, c->d, c->az, c->au, c->av, H, hn, hun, hvn, buoy, zz, idpdx,
                        idpdy, c->dx, c->dy));
  return GT_OK;
}

extern "C" int gt_ip_shchepetkin_mcwilliams(gt_ctx* c, const double* H, const double* hn,
                                            const double* hun, const double* hvn,
                                            const double* buoy, const double* sseo, double* zz,
                                            double* idpdx, double* idpdy) {
  if (!c) return set_err(GT_EARG, "gt_ip_shchepetkin_mcwilliams: NULL context");
  if (c->kmax < 2) return set_err(GT_EARG, "ip_shchepetkin_mcwilliams: kmax must be >= 2");
  ENSURE(c->P, c->N3);
  LCHK(gt::launch_ip_smcw(c->st, c->d, c->az, c->au, c->av, H, hn, hun, hvn, buoy, sseo, zz,
                          c->P, idpdx, idpdy, c->dx, c->dy));
  return GT_OK;
}

static int eos_ip(gt_ctx* c, const char* who, int eqstate_method, int nonnegsalt, int methods,
                  const double* T, const double* S, const double* rho, double T0, double S0,
                  double dtr0, double dsr0, const double* H, const double* hn,
                  const double* hun, const double* hvn, const double* sseo, double* zz_bm,
                  double* idpdx_bm, double* idpdy_bm, double* zz_sm, double* idpdx_sm,
                  double* idpdy_sm) {
  if (!c) return set_err(GT_EARG, "%s: NULL context", who);
  if (eqstate_method < 1 || eqstate_method > 3)
    return set_err(GT_EARG, "init_eqstate(): not a valid eqstate_method");
  if (methods < 1 || methods > 3)
    return set_err(GT_EARG, "%s: methods must be 1 (Blumberg-Mellor), 2 (Shchepetkin-McWilliams) "
                            "or 3 (both)", who);
  if (c->kmax < 2) return set_err(GT_EARG, "%s: kmax must be >= 2", who);
  const bool bm = (methods & 1) != 0, sm = (methods & 2) != 0;
  if (!T || !S || !H || !hn || !hun || !hvn || (eqstate_method != 1 && !rho) ||
      (bm && (!zz_bm || !idpdx_bm || !idpdy_bm)) ||
      (sm && (!zz_sm || !idpdx_sm || !idpdy_sm || !sseo)))
    return set_err(GT_EARG, "%s: NULL argument", who);
  LCHK(gt::launch_eos_ip_fused(c->st, c->d, eqstate_method, nonnegsalt != 0, bm, sm, c->az,
                               c->au, c->av, T, S, rho, T0, S0, dtr0, dsr0, H, hn, hun, hvn,
                               sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm, c->dx,
                               c->dy));
  return GT_OK;
}

extern "C" int gt_eqstate_internal_pressure(
    gt_ctx* c, int eqstate_method, int nonnegsalt, int methods, const double* T, const double* S,
    const double* rho, double T0, double S0, double dtr0, double dsr0, const double* H,
    const double* hn, const double* hun, const double* hvn, const double* sseo, double* zz_bm,
    double* idpdx_bm, double* idpdy_bm, double* zz_sm, double* idpdx_sm, double* idpdy_sm) {
  return eos_ip(c, "gt_eqstate_internal_pressure", eqstate_method, nonnegsalt, methods, T, S,
                rho, T0, S0, dtr0, dsr0, H, hn, hun, hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm,
                idpdx_sm, idpdy_sm);
}

extern "C" int gt_eos_internal_pressure(gt_ctx* c, int methods, const double* T, const double* S,
                                        double T0, double S0, double dtr0, double dsr0,
                                        const double* H, const double* hn, const double* hun,
                                        const double* hvn, const double* sseo, double* zz_bm,
                                        double* idpdx_bm, double* idpdy_bm, double* zz_sm,
                                        double* idpdx_sm, double* idpdy_sm) {
  return eos_ip(c, "gt_eos_internal_pressure", 1, 0, methods, T, S, nullptr, T0, S0, dtr0, dsr0,
                H, hn, hun, hvn, sseo, zz_bm, idpdx_bm, idpdy_bm, zz_sm, idpdx_sm, idpdy_sm);
}

extern "C" int gt_do_internal_pressure(gt_ctx* c, int ip_method, const double* H,
                                       const double* hn, const double* hun, const double* hvn,
                                       const double* buoy, const double* sseo, double* idpdx,
                                       double* idpdy) {
  if (!c) return set_err(GT_EARG, "gt_do_internal_pressure: NULL context");
  switch (ip_method) {
    case GT_IP_BLUMBERG_MELLOR:
      TRY(gt_ip_blumberg_mellor(c, H, hn, hun, hvn, buoy, c->zz, idpdx, idpdy));
      break;
    case GT_IP_SHCHEPETKIN_MCWILLIAMS:
      TRY(gt_ip_shchepetkin_mcwilliams(c, H, hn, hun, hvn, buoy, sseo, c->zz, idpdx, idpdy));
      break;
    case 2: case 3: case 4: case 5: case 7:
      return set_err(GT_ENOTIMPL, "do_internal_pressure: ip_method %d is not part of this build",
                     ip_method);
    default:
      return set_err(GT_EINVAL_SCHEME, "Not valid ip_method specified");
  }
  if (c->sip) {
    LCHK(gt::launch_substr_ini_press(c->st, c->d, c->sip_first, idpdx, idpdy, c->idp0,
                                     c->idp0 + c->N3));
    c->sip_first = 0;
  }
  return GT_OK;
}

extern "C" int gt_set_substr_ini_press(gt_ctx* c, int on) {
  if (!c) return set_err(GT_EARG, "gt_set_substr_ini_press: NULL context");
  if (on && !c->idp0) {
    HIPCHK(hipMalloc(&c->idp0, sizeof(double) * 2 * (size_t)c->N3));
    HIPCHK(hipMemsetAsync(c->idp0, 0, sizeof(double) * 2 * (size_t)c->N3, c->st));
  }
  c->sip = on != 0;
  c->sip_first = 1;
  return GT_OK;
}

// ---- ghost cells ---------------------------------------------------------------
extern "C" int gt_update_3d_halo(gt_ctx* c, double* f1, const double* f2, int tag, int mirror) {
  (void)tag;
  if (!c) return set_err(GT_EARG, "gt_update_3d_halo: NULL context");
  if (mirror != 1) return GT_OK;
  LCHK(gt::launch_mirror_3d(c->st, c->d, f1, f2));
  return GT_OK;
}

extern "C" int gt_update_2d_halo(gt_ctx* c, double* f1, const double* f2, int tag, int mirror) {
  (void)tag;
  if (!c) return set_err(GT_EARG, "gt_update_2d_halo: NULL context");
  if (mirror >= 0) c->do_mirror_2d = mirror != 0;
  if (!c->do_mirror_2d) return GT_OK;
  LCHK(gt::launch_mirror_2d(c->st, c->d, f1, f2));
  return GT_OK;
}

extern "C" int gt_wait_halo(gt_ctx* c, int tag) {
  (void)tag;
  return gt_synchronize(c);
}

extern "C" int gt_exchange_tiles(int px, int py, const int* ni, const int* nj, int nk,
                                 double* const* tiles, void* stream) {
  if (px < 1 || py < 1 || !ni || !nj || !tiles || nk < 1)
    return set_err(GT_EARG, "gt_exchange_tiles: bad argument");
  const int nt = px * py;
  for (int ty = 0; ty < py; ty++)
    for (int tx = 0; tx < px; tx++) {
      int t = ty * px + tx;
      if (ni[t] < GT_HALO || nj[t] < GT_HALO || !tiles[t])
        return set_err(GT_EARG, "gt_exchange_tiles: tile %d smaller than the halo", t);
      if (nj[t] != nj[ty * px] || ni[t] != ni[tx])
        return set_err(GT_EARG, "gt_exchange_tiles: tiles of one row/column differ in size");
    }
  if (nt > gt::kMaxTiles)
    return set_err(GT_EARG, "gt_exchange_tiles: at most %d tiles (got %d)", gt::kMaxTiles, nt);
  // the tile table travels in the kernel argument: nothing is allocated and
  // the call returns once the exchange is enqueued
  gt::TileSet ts{};
  ts.ring_off[0] = 0;
  for (int t = 0; t < nt; t++) {
    ts.t[t] = {tiles[t], ni[t], nj[t]};
    long NX = ni[t] + 2 * GT_HALO;
    ts.ring_off[t + 1] = ts.ring_off[t] + 2 * GT_HALO * NX + 2 * GT_HALO * (long)nj[t];
  }
  LCHK(gt::launch_tile_exchange((hipStream_t)stream, nt, px, ts, nk));
  return GT_OK;
}

// ---- batched multi-tracer step (extension) ----------------------------------
extern "C" int gt_set_tuning(gt_ctx* c, int fused, int tile, int group) {
  if (!c) return set_err(GT_EARG, "gt_set_tuning: NULL context");
  if (tile < 0 || tile > 11)
    return set_err(GT_EARG, "gt_set_tuning: tile must be 0..11 (rows per strip, see the header)");
  if (group != 1 && group != 2 && group != 3 && group != 6 && group != 9)
    return set_err(GT_EARG, "gt_set_tuning: group must be 1, 2, 3, 6 or 9");
  c->fused = fused != 0;
  c->tile = tile;
  c->group = group;
  return GT_OK;
}

// sizes every work field of the transport step up front, so later calls
// neither allocate nor synchronise the device to grow one
static int ensure_bbatch(gt_ctx* c, int ntr);

extern "C" int gt_reserve(gt_ctx* c, int ntracers_f32, int ntracers_f64) {
  if (!c || ntracers_f32 < 0 || ntracers_f64 < 0)
    return set_err(GT_EARG, "gt_reserve: bad argument");
  // the batched step's work fields with room for the budgets formed in it
  size_t ws = strip_ws_bytes(c, sizeof(double), 1, true);
  const size_t w32 = strip_ws_bytes(c, sizeof(float), ntracers_f32, false) +
                     (ntracers_f32 > 0 ? budget_ws_bytes(c, ntracers_f32) : 0);
  const size_t w64 = strip_ws_bytes(c, sizeof(double), ntracers_f64, false) +
                     (ntracers_f64 > 0 ? budget_ws_bytes(c, ntracers_f64) : 0);
  if (w32 > ws) ws = w32;
  if (w64 > ws) ws = w64;
  TRY(ensure_ws(c, ws));
  const long N3 = c->N3;
  ENSURE(c->fi, N3);
  ENSURE(c->hi, N3);
  ENSURE(c->adv, N3);
  ENSURE(c->tmp, N3);
  ENSURE(c->P, N3);
  ENSURE(c->slice3, 3 * c->d.NXY);
  ENSURE(c->bpart, 4 * 1024);
  ENSURE(c->bout, 2);
  if (ntracers_f32 > 0) TRY(ensure_bbatch(c, ntracers_f32));
  if (c->adv_ver_iterations > 1) ENSURE(c->scr7, 7 * N3);
  return GT_OK;
}

extern "C" int gt_set_adjoint_mode(gt_ctx* c, int keep_states) {
  if (!c) return set_err(GT_EARG, "gt_set_adjoint_mode: NULL context");
  if (keep_states < -1 || keep_states > 1)
    return set_err(GT_EARG, "gt_set_adjoint_mode: keep_states must be -1, 0 or 1");
  c->adj_keep = keep_states;
  return GT_OK;
}

// ---- the FP32 step one sub-step at a time ---------------------------------------
static int ensure_w32(gt_ctx* c, size_t bytes) {
  if (c->w32_bytes >= bytes) return GT_OK;
  if (c->w32) {
    HIPCHK(hipDeviceSynchronize());  // the old fields may still be in use on any stream
    HIPCHK(hipFree(c->w32));
    c->w32 = nullptr;
    c->w32_bytes = 0;
  }
  HIPCHK(hipMalloc(&c->w32, bytes));
  c->w32_bytes = bytes;
  return GT_OK;
}


// the five sub-steps of one HALFSPLIT step (advection_3d.F90:439-547) for
// ntr tracers, one kernel each: in -> s[0] (u/2) -> s[1] (v/2) -> s[2] (w)
// -> s[3] (v/2) -> out (u/2, skipped if out is NULL).  D[0..5]: the layer
// heights before/after each sub-step.  Consecutive buffers must differ (s
// may repeat two buffers).
template <typename TF>
static int substeps(gt_ctx* c, hipStream_t st, const GridMasks* g, const double* const D[6],
                    double dt, int ntr, const TF* in, TF* const s[4], TF* out, const double* uu,
                    const double* vv, const double* ww, const double* hu, const double* hv, int hs,
                    int vs, double AH, int kshift) {
  const long ts = c->N3;
  const Dims d = c->d;
  LCHK(gt::launch_sub_h<TF>(st, d, 0, ntr, ts, in, s[0], D[0], D[1], uu, hu, g->flags, dt, c->dx,
                            c->dy, c->ard1, hs, AH));
  LCHK(gt::launch_sub_h<TF>(st, d, 1, ntr, ts, s[0], s[1], D[1], D[2], vv, hv, g->flags, dt,
                            c->dx, c->dy, c->ard1, hs, AH));
  LCHK(gt::launch_sub_w<TF>(st, d, ntr, ts, s[1], s[2], D[2], D[3], ww, g->flags, dt, vs, kshift));
  LCHK(gt::launch_sub_h<TF>(st, d, 1, ntr, ts, s[2], s[3], D[3], D[4], vv, hv, g->flags, dt,
                            c->dx, c->dy, c->ard1, hs, AH));
  if (out)
    LCHK(gt::launch_sub_h<TF>(st, d, 0, ntr, ts, s[3], out, D[4], D[5], uu, hu, g->flags, dt,
                              c->dx, c->dy, c->ard1, hs, AH));
  return GT_OK;
}

// gt_adv3d_batch_f32 with the per-sub-step kernels (tuning fused = 0), one
// domain: work = D1..D5 + two tracer sets
static int batch_substeps_f32(const gt_batch_domain& bd, const GridMasks* g, int ntr, double dt,
                              int hs, int vs, double AH, int kshift, hipStream_t st) {
  gt_ctx* c = bd.ctx;
  const long N3 = c->N3;
  const size_t set = (size_t)ntr * N3;
  TRY(ensure_w32(c, sizeof(double) * 5 * (size_t)N3 + sizeof(float) * 2 * set));
  double* Dh[5];
  for (int n = 0; n < 5; n++) Dh[n] = (double*)c->w32 + n * N3;
  float* A = (float*)((double*)c->w32 + 5 * N3);
  float* B = A + set;
  LCHK(gt::launch_heights(st, c->d, g->flags, bd.ho, bd.uu, bd.vv, bd.ww, Dh, dt, c->dx, c->dy,
                          c->ard1, hs, kshift));
  const double* D[6] = {bd.ho, Dh[0], Dh[1], Dh[2], Dh[3], Dh[4]};
  float* const s[4] = {A, B, A, B};
  TRY(substeps<float>(c, st, g, D, dt, ntr, (const float*)bd.f_in, s, (float*)bd.f_out, bd.uu,
                       bd.vv, bd.ww, bd.hu, bd.hv, hs, vs, AH, kshift));
  if (bd.hires)
    HIPCHK(hipMemcpyAsync(bd.hires, Dh[4], sizeof(double) * N3, hipMemcpyDeviceToDevice, st));
  if (bd.budget) {
    TRY(ensure_bbatch(c, ntr));
    LCHK(gt::launch_budget_batch<float>(st, c->d, c->az, ntr, N3, Dh[4], (const float*)bd.f_out,
                                        c->bbatch, kBudgetBlocks, c->dx * c->dy, bd.budget));
  }
  return GT_OK;
}

template <typename TF>
static int batch_impl(int ndom, const gt_batch_domain* doms, int ntracers, double dt,
                      int hscheme, int vscheme, double AH, int tag, void* stream) {
  if (ndom < 1 || !doms || ntracers < 1)
    return set_err(GT_EARG, "gt_adv3d_batch: bad argument");
  if (!split_scheme(hscheme) || !split_scheme(vscheme))
    return set_err(GT_EINVAL_SCHEME, "gt_adv3d_batch: schemes must be UPSTREAM..P2_PDM");
  // FP32 with fused = 0 on the first context: the per-sub-step kernels
  const bool sub = sizeof(TF) == 4 && doms[0].ctx && !doms[0].ctx->fused;
  if (!sub && !gt::fused_supported(hscheme, vscheme))
    return set_err(GT_ENOTIMPL, "gt_adv3d_batch: hscheme and vscheme must be equal");
  const GridMasks* gsel = nullptr;
  for (int i = 0; i < ndom; i++) {
    gt_ctx* c = doms[i].ctx;
    if (!c || !doms[i].f_in || !doms[i].f_out || doms[i].f_in == doms[i].f_out)
      return set_err(GT_EARG, "gt_adv3d_batch: domain %d needs a context and distinct "
                              "f_in/f_out", i);
    if (c->kmax < 2) return set_err(GT_EARG, "gt_adv3d_batch: kmax must be >= 2");
    if (!strip_fits(c))
      return set_err(GT_ENOTIMPL, "gt_adv3d_batch: domain %d has %ld cells; one field must "
                                  "stay below 4 GiB in FP64 (split the domain)", i, c->N3);
    gsel = grid_for(c, tag);
    if (!gsel) return set_err(GT_EINVAL_TAG, "gt_adv3d_batch: tag is invalid");
    if (doms[i].budget && gsel != &c->H)
      return set_err(GT_EARG, "gt_adv3d_batch: budgets are formed on the T points (tags H, D, "
                              "W); domain %d asks for one on a U or V grid", i);
    // the strip kernels keep their work fields in the context: one context
    // twice in a launch would have two domains write the same work fields
    for (int q = 0; q < i; q++)
      if (doms[q].ctx == c)
        return set_err(GT_EARG, "gt_adv3d_batch: domains %d and %d share one context (give "
                                "each domain its own)", q, i);
  }
  gt_ctx* c0 = doms[0].ctx;
  hipStream_t st = stream ? (hipStream_t)stream : c0->st;
  if (sub) {
    for (int i = 0; i < ndom; i++)
      TRY(batch_substeps_f32(doms[i], grid_for(doms[i].ctx, tag), ntracers, dt, hscheme, vscheme,
                             AH, tag == GT_W_TAG ? 1 : 0, st));
    return GT_OK;
  }
  // FP64: two tracers per wavefront (T and S) when there are two or more,
  // unless the tuning asks for one
  const int group = sizeof(TF) == 4 ? c0->group : (ntracers >= 2 && c0->group != 1 ? 2 : 1);
  for (int i = 0; i < ndom; i++) {
    gt_ctx* c = doms[i].ctx;
    TRY(ensure_ws(c, strip_ws_bytes(c, sizeof(TF), ntracers, false) +
                         (doms[i].budget ? budget_ws_bytes(c, ntracers) : 0)));
  }
  for (int b = 0; b < ndom; b += gt::kMaxFusedDomains) {
    gt::FusedArgs<TF> A{};
    int nd = ndom - b < gt::kMaxFusedDomains ? ndom - b : gt::kMaxFusedDomains;
    for (int i = 0; i < nd; i++) {
      const gt_batch_domain& bd = doms[b + i];
      gt_ctx* c = bd.ctx;
      fill_domain<TF>(c, grid_for(c, tag), &A.doms[i], (const TF*)bd.f_in, (TF*)bd.f_out,
                      c->N3, ntracers, bd.uu, bd.vv, bd.ww, bd.hu, bd.hv, bd.ho, bd.hires,
                      nullptr);
      // the budget partials follow the work fields
      if (bd.budget)
        A.doms[i].bpart = (double*)((char*)c->ws + strip_ws_bytes(c, sizeof(TF), ntracers, false));
    }
    A.ndom = nd;
    A.ntracers = ntracers;
    gt::strip_rows(c0->tile, A.rows);
    A.dt = dt; A.AH = AH;
    A.hscheme = hscheme; A.vscheme = vscheme; A.kshift = tag == GT_W_TAG ? 1 : 0;
    gt::strip_plan(A, group);
    if constexpr (sizeof(TF) == 4) {
      LCHK(gt::launch_fused_f32(st, A, group));
    } else {
      LCHK(gt::launch_fused_f64(st, A, group, false));
    }
    // the domains' budgets from their strip-level partials, one launch
    gt::BudgetPairsSet S{};
    for (int i = 0; i < nd; i++) {
      const gt::FusedDomain<TF>& D = A.doms[i];
      if (!D.bpart) continue;
      S.d[S.n++] = {D.bpart, doms[b + i].budget, (long)D.nsx * D.nsy[1] * D.d.kmax,
                    D.dx * D.dy, ntracers + 1};
    }
    LCHK(gt::launch_budget_pairs_final(st, S));
  }
  return GT_OK;
}

extern "C" int gt_adv3d_batch_f32(int ndom, const gt_batch_domain* doms, int ntracers,
                                  double dt, int hscheme, int vscheme, double AH, int tag,
                                  void* stream) {
  return batch_impl<float>(ndom, doms, ntracers, dt, hscheme, vscheme, AH, tag, stream);
}

extern "C" int gt_adv3d_batch_f64(int ndom, const gt_batch_domain* doms, int ntracers,
                                  double dt, int hscheme, int vscheme, double AH, int tag,
                                  void* stream) {
  return batch_impl<double>(ndom, doms, ntracers, dt, hscheme, vscheme, AH, tag, stream);
}

// ---- autotuner (extension) ------------------------------------------------------
// Times the batched step for every rows-per-strip x tracers-per-wavefront pair
// (reps launches each, after one warm-up) with HIP events on the stream and
// keeps the fastest on every context of the batch.  Outputs are written to
// f_out (f_in is not modified), so the sweep can run on live buffers.
extern "C" int gt_autotune_batch(int ndom, const gt_batch_domain* doms, int ntracers, double dt,
                                 int hscheme, int vscheme, double AH, int tag, int f32, int reps,
                                 void* stream, int* best_tile, int* best_group, float* best_ms,
                                 float* table) {
  if (ndom < 1 || !doms || !doms[0].ctx || reps < 1)
    return set_err(GT_EARG, "gt_autotune_batch: bad argument");
  hipStream_t st = stream ? (hipStream_t)stream : doms[0].ctx->st;
  // strip heights of the two kernels: 9 pairs {16, 32, 8} x {16, 32, 8}
  const int tiles[9] = {3, 4, 5, 6, 7, 8, 9, 10, 11};
  const int groups[4] = {1, 3, 6, 9};
  const int groups64[2] = {1, 2};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 1e30f;
  int bt = 0, bg = 3;
  // two passes over the candidates, each candidate's time the lower of its
  // two: a transient slowdown of the device then moves no choice
  float tmin_ms[36];
  for (int q = 0; q < 36; q++) tmin_ms[q] = 1e30f;
  for (int pass = 0; pass < 2; pass++) {
  int n = 0;
  for (int ti = 0; ti < 9; ti++)
    for (int gi = 0; gi < (f32 ? 4 : 2); gi++) {
      const int grp = f32 ? groups[gi] : groups64[gi];
      for (int i = 0; i < ndom; i++) {
        doms[i].ctx->tile = tiles[ti];
        doms[i].ctx->group = grp;
      }
      auto run = [&]() -> int {
        return f32 ? gt_adv3d_batch_f32(ndom, doms, ntracers, dt, hscheme, vscheme, AH, tag, st)
                   : gt_adv3d_batch_f64(ndom, doms, ntracers, dt, hscheme, vscheme, AH, tag, st);
      };
      int rc = run();
      if (rc != GT_OK) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
      }
      HIPCHK(hipEventRecord(e0, st));
      for (int r = 0; r < reps; r++) TRY(run());
      HIPCHK(hipEventRecord(e1, st));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      ms /= reps;
      ms = tmin_ms[n] = ms < tmin_ms[n] ? ms : tmin_ms[n];
      if (table) table[n] = ms;
      n++;
      if (pass == 1 && ms < best) {
        best = ms;
        bt = tiles[ti];
        bg = grp;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int i = 0; i < ndom; i++) {
    doms[i].ctx->tile = bt;
    doms[i].ctx->group = bg;
  }
  if (best_tile) *best_tile = bt;
  if (best_group) *best_group = bg;
  if (best_ms) *best_ms = best;
  return GT_OK;
}

// ---- reverse mode (extension) ------------------------------------------------
// transpose of one HALFSPLIT step from s0: the states before the last four
// sub-steps are in S[0..3] (have_states) or are recomputed there; lam[0]
// (adjoint of the step's output) is replaced by the adjoint of s0, lam[1] is
// work
template <typename TF>
static int reverse_step(gt_ctx* c, const GridMasks* g, const double* const D[6], double dt,
                        int ntr, const TF* s0, TF* const S[4], TF** lam, const double* uu,
                        const double* vv, const double* ww, const double* hu, const double* hv,
                        int hs, int vs, double AH, int kshift, bool have_states) {
  hipStream_t st = c->st;
  const Dims d = c->d;
  const long ts = c->N3;
  if (!have_states)
    TRY(substeps<TF>(c, st, g, D, dt, ntr, s0, S, nullptr, uu, vv, ww, hu, hv, hs, vs, AH,
                     kshift));
  TF* a = lam[0];
  TF* b = lam[1];
  LCHK(gt::launch_sub_h_adj<TF>(st, d, 0, ntr, ts, S[3], a, b, D[4], D[5], uu, hu, g->flags, dt,
                                c->dx, c->dy, c->ard1, hs, AH));
  LCHK(gt::launch_sub_h_adj<TF>(st, d, 1, ntr, ts, S[2], b, a, D[3], D[4], vv, hv, g->flags, dt,
                                c->dx, c->dy, c->ard1, hs, AH));
  LCHK(gt::launch_sub_w_adj<TF>(st, d, ntr, ts, S[1], a, b, D[2], D[3], ww, g->flags, dt, vs,
                                kshift));
  LCHK(gt::launch_sub_h_adj<TF>(st, d, 1, ntr, ts, S[0], b, a, D[1], D[2], vv, hv, g->flags, dt,
                                c->dx, c->dy, c->ard1, hs, AH));
  LCHK(gt::launch_sub_h_adj<TF>(st, d, 0, ntr, ts, s0, a, b, D[0], D[1], uu, hu, g->flags, dt,
                                c->dx, c->dy, c->ard1, hs, AH));
  lam[0] = b;
  lam[1] = a;
  return GT_OK;
}

template <typename TF>
static int copy_set(hipStream_t st, TF* dst, const TF* src, size_t n) {
  if (dst == src || n == 0) return GT_OK;
  HIPCHK(hipMemcpyAsync(dst, src, sizeof(TF) * n, hipMemcpyDeviceToDevice, st));
  return GT_OK;
}

// the inputs of each step: one set for every step (velocities and layer
// heights held fixed) or one set per step
struct StepInputs {
  const double* const* f[6];  // uu vv ww hu hv ho, indexed by step (or [0] for all)
  bool per_step;
  const double* get(int n, int t) const { return f[n][per_step ? t : 0]; }
};

// Forward sweep from f0 with a checkpoint every ckpt_every steps (the strip
// kernels where they apply, else the per-sub-step kernels: the same states
// bit for bit), then the reverse sweep segment by segment: each segment's
// states are recomputed from its checkpoint, then every step is transposed
// from the last.  The layer heights of the five sub-steps depend on the
// step's inputs only: formed once when they are held fixed, per step when
// they change.
template <typename TF>
static int adjoint_impl(const char* name, gt_ctx* c, int ntr, int nsteps, int ckpt_every,
                        double dt, const TF* f0, TF* f_final, const TF* lam_final, TF* lam0,
                        const StepInputs& in, int hscheme, int vscheme, double AH, int tag) {
  if (!c || ntr < 1 || !f0 || !lam_final || !lam0 || nsteps < 1 || ckpt_every < 1)
    return set_err(GT_EARG, "%s: bad argument", name);
  for (int n = 0; n < 6; n++) {
    if (!in.f[n]) return set_err(GT_EARG, "%s: NULL velocity or layer height", name);
    for (int t = 0; t < (in.per_step ? nsteps : 1); t++)
      if (!in.f[n][t]) return set_err(GT_EARG, "%s: NULL velocity or layer height", name);
  }
  const GridMasks* g = grid_for(c, tag);
  if (!g) return set_err(GT_EINVAL_TAG, "%s: tag is invalid", name);
  if (!split_scheme(hscheme) || !split_scheme(vscheme))
    return set_err(GT_EINVAL_SCHEME, "%s: schemes must be UPSTREAM..P2_PDM", name);
  if (c->kmax < 2) return set_err(GT_EARG, "%s: kmax must be >= 2", name);
  const int kshift = tag == GT_W_TAG ? 1 : 0;
  const long N3 = c->N3;
  const size_t set = (size_t)ntr * N3;
  const int nck = (nsteps + ckpt_every - 1) / ckpt_every;
  const int seg = ckpt_every < nsteps ? ckpt_every : nsteps;  // steps in a segment, at most
  // Work memory: the layer heights D1..D5 (one slot, or one per step of a
  // segment when the states are kept and the inputs change), then tracer
  // sets: the sub-step states (4, or 4 per step of a segment when kept),
  // lam[2], the checkpoints and the seg-1 states inside a segment.  Keeping
  // the states lets the reverse steps skip the four sub-steps that would
  // recover them; it is taken when asked for, or (default) when it needs at
  // most half of the free device memory.
  const size_t rest = 2 + (size_t)nck + (size_t)(seg - 1);
  auto bytes_for = [&](bool k) -> size_t {
    const size_t hs = k && in.per_step ? (size_t)seg : 1;
    return sizeof(double) * 5 * (size_t)N3 * hs + sizeof(TF) * ((k ? 4 * (size_t)seg : 4) + rest) * set;
  };
  bool keep = c->adj_keep == 1;
  if (c->adj_keep < 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) keep = bytes_for(true) <= (fr + c->w32_bytes) / 2;
  }
  int got = ensure_w32(c, bytes_for(keep));
  if (got != GT_OK && keep && c->adj_keep < 0) {
    // the free memory was there but not in one piece: recompute instead
    (void)hipGetLastError();
    keep = false;
    got = ensure_w32(c, bytes_for(false));
  }
  if (got != GT_OK) return got;
  const int hslots = keep && in.per_step ? seg : 1;
  double* hbase = (double*)c->w32;
  TF* base = (TF*)(hbase + 5 * (size_t)N3 * hslots);
  const int nS = keep ? 4 * seg : 4;
  TF* lam[2] = {base + (size_t)nS * set, base + (size_t)(nS + 1) * set};
  TF* ck = base + (size_t)(nS + 2) * set;
  TF* sg = ck + (size_t)nck * set;
  auto sub_states = [&](int n, TF* S[4]) {  // the 4 sub-step states of set group n
    for (int q = 0; q < 4; q++) S[q] = base + (size_t)(4 * n + q) * set;
  };
  std::vector<int> slot_key(hslots, -1);  // the step whose heights a slot holds
  auto heights = [&](int t, int slot, const double* D[6]) -> int {
    double* Hs = hbase + 5 * (size_t)N3 * slot;
    double* Dh[5];
    for (int n = 0; n < 5; n++) Dh[n] = Hs + n * N3;
    const int key = in.per_step ? t : 0;
    if (slot_key[slot] != key) {
      LCHK(gt::launch_heights(c->st, c->d, g->flags, in.get(5, t), in.get(0, t), in.get(1, t),
                              in.get(2, t), Dh, dt, c->dx, c->dy, c->ard1, hscheme, kshift));
      slot_key[slot] = key;
    }
    D[0] = in.get(5, t);
    for (int n = 0; n < 5; n++) D[n + 1] = Dh[n];
    return GT_OK;
  };
  const bool plain = c->fused && gt::fused_supported(hscheme, vscheme) && strip_fits(c);
  // one step from src to dst; S: where the sub-step kernels leave the states
  // (the strip kernels where they apply and the states are not wanted)
  auto step = [&](int t, int slot, const TF* src, TF* dst, TF* const S[4], bool want) -> int {
    if (plain && !want) {
      gt_batch_domain bd{c, src, dst, in.get(0, t), in.get(1, t), in.get(2, t), in.get(3, t),
                         in.get(4, t), in.get(5, t), nullptr};
      return batch_impl<TF>(1, &bd, ntr, dt, hscheme, vscheme, AH, tag, c->st);
    }
    const double* D[6];
    TRY(heights(t, slot, D));
    TF* const s[4] = {S[0], S[1], want ? S[2] : S[0], want ? S[3] : S[1]};
    return substeps<TF>(c, c->st, g, D, dt, ntr, src, s, dst, in.get(0, t), in.get(1, t),
                        in.get(2, t), in.get(3, t), in.get(4, t), hscheme, vscheme, AH, kshift);
  };
  TF* S0[4];
  sub_states(0, S0);
  // forward sweep (lam[0], lam[1] as the two states)
  TF* a = lam[0];
  TF* b = lam[1];
  TRY(copy_set<TF>(c->st, a, f0, set));
  // With the states kept, the last segment keeps them already here (its
  // states inside the segment go where its recompute would put them), so
  // the reverse sweep need not recompute it.
  const int tl = (nck - 1) * ckpt_every;  // the last segment's first step
  for (int t = 0; t < nsteps; t++) {
    if (t % ckpt_every == 0) TRY(copy_set<TF>(c->st, ck + (size_t)(t / ckpt_every) * set, a, set));
    if (keep && t >= tl) {
      TF* St[4];
      sub_states(t - tl, St);
      const TF* src = t == tl ? a : sg + (size_t)(t - tl - 1) * set;
      TF* dst = t < nsteps - 1 ? sg + (size_t)(t - tl) * set : b;
      TRY(step(t, hslots > 1 ? t - tl : 0, src, dst, St, true));
      if (t == nsteps - 1) {
        TF* x = a; a = b; b = x;
      }
      continue;
    }
    TRY(step(t, 0, a, b, S0, false));
    TF* x = a; a = b; b = x;
  }
  if (f_final) TRY(copy_set<TF>(c->st, f_final, a, set));
  // reverse sweep, segment by segment from the last checkpoint
  TRY(copy_set<TF>(c->st, lam[0], lam_final, set));
  for (int s = nck - 1; s >= 0; s--) {
    const int t0 = s * ckpt_every, t1 = t0 + ckpt_every < nsteps ? t0 + ckpt_every : nsteps;
    auto state = [&](int t) -> TF* {
      return t == t0 ? ck + (size_t)s * set : sg + (size_t)(t - t0 - 1) * set;
    };
    auto slot = [&](int t) -> int { return hslots > 1 ? t - t0 : 0; };
    if (keep && s < nck - 1) {
      // recompute the segment through the sub-step kernels, keeping every
      // step's four states (the last step's output is not needed)
      for (int t = t0; t < t1; t++) {
        TF* St[4];
        sub_states(t - t0, St);
        TRY(step(t, slot(t), state(t), t < t1 - 1 ? state(t + 1) : nullptr, St, true));
      }
    } else if (!keep) {
      for (int t = t0; t < t1 - 1; t++) TRY(step(t, 0, state(t), state(t + 1), S0, false));
    }
    for (int t = t1 - 1; t >= t0; t--) {
      const double* D[6];
      TRY(heights(t, slot(t), D));
      TF* St[4];
      sub_states(keep ? t - t0 : 0, St);
      TRY(reverse_step<TF>(c, g, D, dt, ntr, state(t), St, lam, in.get(0, t), in.get(1, t),
                           in.get(2, t), in.get(3, t), in.get(4, t), hscheme, vscheme, AH,
                           kshift, keep));
    }
  }
  return copy_set<TF>(c->st, lam0, lam[0], set);
}

// one set of inputs for every step
struct FixedInputs {
  const double* p[6];
  StepInputs in;
  FixedInputs(const double* uu, const double* vv, const double* ww, const double* hu,
              const double* hv, const double* ho)
      : p{uu, vv, ww, hu, hv, ho} {
    for (int n = 0; n < 6; n++) in.f[n] = &p[n];
    in.per_step = false;
  }
};

extern "C" int gt_adv3d_adjoint(gt_ctx* c, int nsteps, int ckpt_every, double dt,
                                const double* f0, double* f_final, const double* lam_final,
                                double* lam0, const double* uu, const double* vv,
                                const double* ww, const double* hu, const double* hv,
                                const double* ho, int hscheme, int vscheme, double AH,
                                int tag) {
  const FixedInputs fi(uu, vv, ww, hu, hv, ho);
  return adjoint_impl<double>("gt_adv3d_adjoint", c, 1, nsteps, ckpt_every, dt, f0, f_final,
                              lam_final, lam0, fi.in, hscheme, vscheme, AH, tag);
}

extern "C" int gt_adv3d_adjoint_f32(gt_ctx* c, int ntracers, int nsteps, int ckpt_every,
                                    double dt, const float* f0, float* f_final,
                                    const float* lam_final, float* lam0, const double* uu,
                                    const double* vv, const double* ww, const double* hu,
                                    const double* hv, const double* ho, int hscheme,
                                    int vscheme, double AH, int tag) {
  const FixedInputs fi(uu, vv, ww, hu, hv, ho);
  return adjoint_impl<float>("gt_adv3d_adjoint_f32", c, ntracers, nsteps, ckpt_every, dt, f0,
                             f_final, lam_final, lam0, fi.in, hscheme, vscheme, AH, tag);
}

extern "C" int gt_adv3d_adjoint_seq(gt_ctx* c, int nsteps, int ckpt_every, double dt,
                                    const double* f0, double* f_final, const double* lam_final,
                                    double* lam0, const double* const* uu,
                                    const double* const* vv, const double* const* ww,
                                    const double* const* hu, const double* const* hv,
                                    const double* const* ho, int hscheme, int vscheme,
                                    double AH, int tag) {
  const StepInputs in{{uu, vv, ww, hu, hv, ho}, true};
  return adjoint_impl<double>("gt_adv3d_adjoint_seq", c, 1, nsteps, ckpt_every, dt, f0,
                              f_final, lam_final, lam0, in, hscheme, vscheme, AH, tag);
}

extern "C" int gt_adv3d_adjoint_seq_f32(gt_ctx* c, int ntracers, int nsteps, int ckpt_every,
                                        double dt, const float* f0, float* f_final,
                                        const float* lam_final, float* lam0,
                                        const double* const* uu, const double* const* vv,
                                        const double* const* ww, const double* const* hu,
                                        const double* const* hv, const double* const* ho,
                                        int hscheme, int vscheme, double AH, int tag) {
  const StepInputs in{{uu, vv, ww, hu, hv, ho}, true};
  return adjoint_impl<float>("gt_adv3d_adjoint_seq_f32", c, ntracers, nsteps, ckpt_every, dt,
                             f0, f_final, lam_final, lam0, in, hscheme, vscheme, AH, tag);
}

extern "C" int gt_ip_blumberg_mellor_adjoint(gt_ctx* c, const double* hun, const double* hvn,
                                             const double* zz, const double* idpdx_bar,
                                             const double* idpdy_bar, double* buoy_bar) {
  if (!c || !hun || !hvn || !zz || !idpdx_bar || !idpdy_bar || !buoy_bar)
    return set_err(GT_EARG, "gt_ip_blumberg_mellor_adjoint: NULL argument");
  if (c->kmax > 64) return set_err(GT_EARG, "gt_ip_blumberg_mellor_adjoint: kmax > 64");
  LCHK(gt::launch_bm_adj(c->st, c->d, c->au, c->av, zz, hun, hvn, idpdx_bar, idpdy_bar,
                         buoy_bar, c->dx, c->dy));
  return GT_OK;
}

extern "C" int gt_ip_shchepetkin_mcwilliams_adjoint(gt_ctx* c, const double* hn,
                                                    const double* hun, const double* hvn,
                                                    const double* buoy, const double* sseo,
                                                    const double* zz, const double* idpdx_bar,
                                                    const double* idpdy_bar, double* buoy_bar) {
  if (!c || !hn || !hun || !hvn || !buoy || !sseo || !zz || !idpdx_bar || !idpdy_bar ||
      !buoy_bar)
    return set_err(GT_EARG, "gt_ip_shchepetkin_mcwilliams_adjoint: NULL argument");
  if (c->kmax < 2 || c->kmax > 64)
    return set_err(GT_EARG, "gt_ip_shchepetkin_mcwilliams_adjoint: kmax must be 2..64");
  ENSURE(c->fi, c->N3);
  ENSURE(c->hi, c->N3);
  LCHK(gt::launch_sm_adj(c->st, c->d, c->az, c->au, c->av, buoy, zz, hn, hun, hvn, sseo,
                         idpdx_bar, idpdy_bar, c->fi, c->hi, buoy_bar, c->dx, c->dy));
  return GT_OK;
}

extern "C" int gt_eos_linear_adjoint(gt_ctx* c, const double* buoy_bar, double* T_bar,
                                     double* S_bar, double dtr0, double dsr0) {
  if (!c || !buoy_bar) return set_err(GT_EARG, "gt_eos_linear_adjoint: NULL argument");
  LCHK(gt::launch_eos_linear_adj(c->st, c->d, buoy_bar, T_bar, S_bar, dtr0, dsr0));
  return GT_OK;
}

// ---- budgets -----------------------------------------------------------------------

template <typename TF>
static int budget_impl(gt_ctx* c, const double* h, const TF* f, double* mass, double* volume) {
  if (!c || !h) return set_err(GT_EARG, "gt_budget: NULL argument");
  const int nblocks = kBudgetBlocks;
  ENSURE(c->bpart, 4 * nblocks);
  ENSURE(c->bout, 2);
  LCHK(gt::launch_budget<TF>(c->st, c->d, c->az, h, f, c->bpart, nblocks, c->dx * c->dy,
                             c->bout));
  double hb[2];
  HIPCHK(hipMemcpyAsync(hb, c->bout, sizeof(hb), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (mass) *mass = hb[0];
  if (volume) *volume = hb[1];
  return GT_OK;
}

extern "C" int gt_budget(gt_ctx* c, const double* h, const double* f, double* mass,
                         double* volume) {
  return budget_impl<double>(c, h, f, mass, volume);
}

extern "C" int gt_budget_f32(gt_ctx* c, const double* h, const float* f, double* mass,
                             double* volume) {
  return budget_impl<float>(c, h, f, mass, volume);
}

static long budget_batch_partials(int ntr) { return 2L * (ntr + 1) * kBudgetBlocks; }

static int ensure_bbatch(gt_ctx* c, int ntr) {
  const long n = budget_batch_partials(ntr);
  if (c->bbatch_n >= n) return GT_OK;
  if (c->bbatch) {
    HIPCHK(hipDeviceSynchronize());  // the old partials may still be in use on any stream
    HIPCHK(hipFree(c->bbatch));
    c->bbatch = nullptr;
    c->bbatch_n = 0;
  }
  HIPCHK(hipMalloc(&c->bbatch, sizeof(double) * (size_t)n));
  c->bbatch_n = n;
  return GT_OK;
}

static int budget_batch_check(const char* name, gt_ctx* c, int ntr, const double* h,
                              const float* f, const double* out) {
  if (!c || !h || !f || !out) return set_err(GT_EARG, "%s: NULL argument", name);
  if (ntr < 1) return set_err(GT_EARG, "%s: ntracers must be >= 1", name);
  return GT_OK;
}

extern "C" int gt_budget_batch_f32(gt_ctx* c, int ntracers, const double* h, const float* f,
                                   double* out) {
  TRY(budget_batch_check("gt_budget_batch_f32", c, ntracers, h, f, out));
  TRY(ensure_bbatch(c, ntracers));
  LCHK(gt::launch_budget_batch<float>(c->st, c->d, c->az, ntracers, c->N3, h, f, c->bbatch,
                                      kBudgetBlocks, c->dx * c->dy, out));
  return GT_OK;
}

// ---- allocation-free budgets --------------------------------------------------------
// gt_budget_batch_f32_noalloc keeps its partial sums in this array instead of
// the context: every GPU the library runs on holds one copy until the process
// ends.  In host code the array's name is a host-side stand-in, not a device
// address: host code uses only the address hipGetSymbolAddress returns for
// the current device, looked up once per device (noalloc_scratch, called by
// gt_create with the context's device current) and kept in the context.
constexpr int kNoallocMaxTracers = 32;
constexpr long kNoallocScratch = 2L * (kNoallocMaxTracers + 1) * kBudgetBlocks;
static __device__ __align__(256) double scratch[kNoallocScratch];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

static int noalloc_scratch(double** p) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_err(GT_EHIP, "gt_create: device %d is beyond the %d the library tracks", dev,
                   kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* q = nullptr;
    HIPCHK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<double*>(q);
  }
  *p = g_scratch_dev[dev];
  return GT_OK;
}

extern "C" int gt_budget_batch_f32_noalloc(gt_ctx* c, int ntracers, const double* h,
                                           const float* f, double* out) {
  TRY(budget_batch_check("gt_budget_batch_f32_noalloc", c, ntracers, h, f, out));
  if (ntracers > kNoallocMaxTracers)
    return set_err(GT_EARG,
                   "gt_budget_batch_f32_noalloc: %d tracers exceed the %d its scratch holds "
                   "(gt_budget_batch_f32 takes any number)",
                   ntracers, kNoallocMaxTracers);
  if (!c->noalloc)
    return set_err(GT_EHIP, "gt_budget_batch_f32_noalloc: no scratch address for this context");
  // every partial the final pass reads is written by the first pass of the
  // same call, so nothing is cleared
  LCHK(gt::launch_budget_batch<float>(c->st, c->d, c->az, ntracers, c->N3, h, f, c->noalloc,
                                      kBudgetBlocks, c->dx * c->dy, out));
  return GT_OK;
}

// ---- memory helpers -----------------------------------------------------------------
extern "C" int gt_malloc(void** p, size_t bytes) {
  if (!p) return set_err(GT_EARG, "gt_malloc: NULL pointer");
  HIPCHK(hipMalloc(p, bytes));
  return GT_OK;
}
extern "C" int gt_free(void* p) {
  if (p) HIPCHK(hipFree(p));
  return GT_OK;
}
extern "C" int gt_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GT_OK;
}
extern "C" int gt_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GT_OK;
}
