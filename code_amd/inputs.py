"""Seeded synthetic inputs for the transport step.

The reference ships no test data for this path, so every workload is built
here from a recipe: bathymetry, land mask, layer heights, transports, vertical
velocity and tracers.  All fields use the reference's array layout
(``f(imin-2:imax+2, jmin-2:jmax+2, 0:kmax)``, i fastest; numpy arrays are
indexed ``[k, j, i]`` with C order so that memory order matches Fortran's).

Recipe sources in getm-model/code:
* land/water mask and u/v/x masks: src/domain/domain.F90:249-346
* HU/HV (vel_depth_method=0): src/domain/uv_depths.F90:74,94
* sigma layers: src/3d/sigma_coordinates.F90:53-104
* general vertical coordinates: src/3d/general_coordinates.F90:56-140
* vertical velocity from continuity: src/3d/ww_momentum_3d.F90:74-106
* linear EOS constants: src/3d/eqstate.F90:26-27
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

HALO = 2
HLAND = -10.0
SMALL = 1e-8

# reference defaults are single-precision literals promoted to FP64
# (src/3d/eqstate.F90:26-27)
EOS_T0 = float(np.float32(10.0))
EOS_S0 = float(np.float32(33.75))
EOS_DTR0 = float(np.float32(-0.17))
EOS_DSR0 = float(np.float32(0.78))


def _splitmix64(n: int, seed: int) -> np.ndarray:
    """n uint64 values of the splitmix64 sequence started at ``seed``."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64)
             * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(shape, seed: int, lo=0.0, hi=1.0) -> np.ndarray:
    n = int(np.prod(shape))
    u = (_splitmix64(n, seed) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return (lo + (hi - lo) * u).reshape(shape)


def normal(shape, seed: int) -> np.ndarray:
    """Box-Muller on two splitmix64 streams."""
    u1 = uniform(shape, seed * 2 + 1)
    u2 = uniform(shape, seed * 2 + 2)
    u1 = np.maximum(u1, 2.0 ** -53)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


@dataclasses.dataclass
class Domain:
    imax: int
    jmax: int
    kmax: int
    dx: float
    dy: float
    H: np.ndarray       # (NY, NX)
    HU: np.ndarray
    HV: np.ndarray
    az: np.ndarray      # int32 (NY, NX)
    au: np.ndarray
    av: np.ndarray
    ax: np.ndarray
    imin: int = 1
    jmin: int = 1

    @property
    def NX(self):
        return self.imax - self.imin + 1 + 2 * HALO

    @property
    def NY(self):
        return self.jmax - self.jmin + 1 + 2 * HALO

    @property
    def shape3(self):
        return (self.kmax + 1, self.NY, self.NX)

    @property
    def ard1(self):
        return 1.0 / (self.dx * self.dy)


def masks_from_az(az: np.ndarray):
    """domain.F90:294-346 (serial build: the mask halo updates are no-ops)."""
    NY, NX = az.shape
    au = np.zeros_like(az)
    av = np.zeros_like(az)
    ax = np.zeros_like(az)
    a, b = az[:, :-1], az[:, 1:]
    u = np.zeros_like(a)
    u[(a == 1) & (b == 1)] = 1
    u[((a == 1) & (b == 2)) | ((a == 2) & (b == 1))] = 2
    u[(a == 2) & (b == 2)] = 3
    au[:, :-1] = u
    a, b = az[:-1, :], az[1:, :]
    v = np.zeros_like(a)
    v[(a == 1) & (b == 1)] = 1
    v[((a == 1) & (b == 2)) | ((a == 2) & (b == 1))] = 2
    v[(a == 2) & (b == 2)] = 3
    av[:-1, :] = v
    x = ((az[:-1, :-1] >= 1) & (az[1:, :-1] >= 1) & (az[:-1, 1:] >= 1)
         & (az[1:, 1:] >= 1))
    ax[:-1, :-1] = x.astype(az.dtype)
    return au, av, ax


def make_domain(imax, jmax, kmax, seed, hmin=10.0, hmax=40.0, walls=5,
                jacobi=5, dx=500.0, dy=500.0) -> Domain:
    """Seeded bathymetry U(hmin,hmax), smoothed by ``jacobi`` 4-neighbour
    passes over the interior, with ``walls`` land cells on every side."""
    NX, NY = imax + 2 * HALO, jmax + 2 * HALO
    H = uniform((NY, NX), seed * 7919 + 11, hmin, hmax)
    for _ in range(jacobi):
        Hn = H.copy()
        Hn[HALO:-HALO, HALO:-HALO] = 0.25 * (
            H[HALO:-HALO, HALO - 1:-HALO - 1] + H[HALO:-HALO, HALO + 1:NX - HALO + 1]
            + H[HALO - 1:-HALO - 1, HALO:-HALO] + H[HALO + 1:NY - HALO + 1, HALO:-HALO])
        H = Hn
    land = np.ones((NY, NX), bool)
    land[HALO + walls:NY - HALO - walls, HALO + walls:NX - HALO - walls] = False
    H[land] = HLAND
    az = np.zeros((NY, NX), np.int32)
    interior = np.zeros((NY, NX), bool)
    interior[HALO:-HALO, HALO:-HALO] = True
    az[interior & (H > HLAND + SMALL)] = 1
    au, av, ax = masks_from_az(az)
    HU = np.zeros_like(H)
    HV = np.zeros_like(H)
    HU[:, :-1] = 0.5 * (H[:, :-1] + H[:, 1:])
    HV[:-1, :] = 0.5 * (H[:-1, :] + H[1:, :])
    return Domain(imax, jmax, kmax, dx, dy, H, HU, HV, az, au, av, ax)


def sigma_layers(dom: Domain, depth2d: np.ndarray) -> np.ndarray:
    """Equidistant sigma: h_k = D/kmax (sigma_coordinates.F90:84-104)."""
    kmaxm1 = 1.0 / float(dom.kmax)
    h = np.zeros(dom.shape3)
    h[1:] = (depth2d * kmaxm1)[None]
    return h


def _gvc_gamma(kmax, D, ddu=1.0, ddl=1.0, d_gamma=20.0):
    """dimensionless gamma-coordinate gga(k), general_coordinates.F90:60-98
    with gamma_surf=.true."""
    k = np.arange(0, kmax + 1, dtype=np.float64)
    be = np.tanh((ddl + ddu) * k / kmax - ddl) + np.tanh(ddl)
    be = be / (np.tanh(ddl) + np.tanh(ddu)) - 1.0
    sig = k / kmax - 1.0
    be[0] = sig[0] = -1.0
    kk = kmax
    alpha = np.minimum(((be[kk] - be[kk - 1]) - d_gamma / D * (sig[kk] - sig[kk - 1]))
                       / ((be[kk] - be[kk - 1]) - (sig[kk] - sig[kk - 1])), 1.0)
    return alpha[None] * sig[:, None, None] + (1.0 - alpha[None]) * be[:, None, None]


def gvc_layers(dom: Domain, depth2d: np.ndarray, min_depth=0.5) -> np.ndarray:
    """GVC layer heights; the top layer closes the column exactly as the
    reference does (hn(kmax) = ssen - zz, general_coordinates.F90:153-160)."""
    D = np.maximum(depth2d, min_depth)
    gga = _gvc_gamma(dom.kmax, D)
    h = np.zeros(dom.shape3)
    h[1:] = D[None] * (gga[1:] - gga[:-1])
    zz = -D + np.sum(h[1:dom.kmax], axis=0)
    h[dom.kmax] = 0.0 - zz
    return h


def streamfunction_transports(dom: Domain, modes, umax):
    """Depth-integrated transports from a streamfunction on X-points that is
    zero on every X-point touching land: U = -dpsi/dy, V = dpsi/dx.  The
    discrete divergence telescopes to zero.  ``modes`` is a list of
    (m, n, amplitude).  Scaled so that max|U/HU|, max|V/HV| = umax."""
    NY, NX = dom.az.shape
    ii = np.arange(NX)[None, :].astype(np.float64)
    jj = np.arange(NY)[:, None].astype(np.float64)
    wat_i = np.where(dom.az.max(axis=0) == 1)[0]
    wat_j = np.where(dom.az.max(axis=1) == 1)[0]
    i0, i1 = wat_i.min(), wat_i.max()
    j0, j1 = wat_j.min(), wat_j.max()
    Lx, Ly = float(i1 - i0 + 1), float(j1 - j0 + 1)
    xp = (ii - (i0 - 1)) / Lx
    yp = (jj - (j0 - 1)) / Ly
    psi = np.zeros((NY, NX))
    for m, n, a in modes:
        psi = psi + a * np.sin(m * math.pi * xp) * np.sin(n * math.pi * yp)
    psi = psi * (dom.ax == 1)
    U = np.zeros((NY, NX))
    V = np.zeros((NY, NX))
    U[1:, :] = -(psi[1:, :] - psi[:-1, :]) / dom.dy
    V[:, 1:] = (psi[:, 1:] - psi[:, :-1]) / dom.dx
    U *= (dom.au >= 1)
    V *= (dom.av >= 1)
    su = np.max(np.abs(U[dom.au >= 1] / dom.HU[dom.au >= 1]))
    sv = np.max(np.abs(V[dom.av >= 1] / dom.HV[dom.av >= 1]))
    s = umax / max(su, sv)
    return U * s, V * s


def layer_velocities(dom: Domain, U, V, hun, hvn):
    """Split depth-integrated transports over layers by h_k/D."""
    Dun = np.sum(hun[1:], axis=0)
    Dvn = np.sum(hvn[1:], axis=0)
    uu = np.zeros(dom.shape3)
    vv = np.zeros(dom.shape3)
    with np.errstate(divide="ignore", invalid="ignore"):
        fu = np.where(dom.au >= 1, U / np.where(Dun == 0, 1, Dun), 0.0)
        fv = np.where(dom.av >= 1, V / np.where(Dvn == 0, 1, Dvn), 0.0)
    uu[1:] = hun[1:] * fu[None]
    vv[1:] = hvn[1:] * fv[None]
    return uu, vv


def ww_continuity(dom: Domain, uu, vv, ho, hn, dt):
    """ww_momentum_3d.F90:74-106 over the interior, ww(0)=ww(kmax)=0."""
    NY, NX = dom.az.shape
    ww = np.zeros(dom.shape3)
    dtm1 = 1.0 / dt
    sl = (slice(HALO, NY - HALO), slice(HALO, NX - HALO))
    slw = (slice(HALO, NY - HALO), slice(HALO - 1, NX - HALO - 1))
    sls = (slice(HALO - 1, NY - HALO - 1), slice(HALO, NX - HALO))
    wet = dom.az[sl] == 1
    for k in range(1, dom.kmax):
        div = (uu[k][sl] * dom.dy - uu[k][slw] * dom.dy
               + vv[k][sl] * dom.dx - vv[k][sls] * dom.dx) * dom.ard1
        val = ww[k - 1][sl] - (hn[k][sl] - ho[k][sl]) * dtm1 - div
        ww[k][sl] = np.where(wet, val, ww[k][sl])
    return ww


def gaussian_blob(dom: Domain, ci, cj, sigma, amp=1.0, background=0.0):
    NY, NX = dom.az.shape
    ii = np.arange(NX)[None, :] - HALO + dom.imin
    jj = np.arange(NY)[:, None] - HALO + dom.jmin
    r2 = (ii - ci) ** 2 + (jj - cj) ** 2
    c = background + amp * np.exp(-r2 / (2.0 * sigma ** 2))
    f = np.zeros(dom.shape3)
    f[1:] = c[None]
    return f


@dataclasses.dataclass
class Transport:
    dom: Domain
    dt: float
    uu: np.ndarray
    vv: np.ndarray
    ww: np.ndarray
    hun: np.ndarray
    hvn: np.ndarray
    ho: np.ndarray
    hn: np.ndarray
    tracers: dict


def config_advection(imax, jmax, kmax, seed, *, hmin=10.0, hmax=40.0,
                     coords="sigma", umax=0.5, modes=None, dt=30.0,
                     dx=500.0, dy=500.0, walls=5) -> Transport:
    """Generic recipe: seeded bathymetry, closed walls, layers, a
    non-divergent streamfunction flow with ww from continuity, steady
    surface (ho == hn)."""
    dom = make_domain(imax, jmax, kmax, seed, hmin, hmax, walls=walls, dx=dx,
                      dy=dy)
    layers = sigma_layers if coords == "sigma" else gvc_layers
    hn = layers(dom, dom.H)
    hun = layers(dom, dom.HU)
    hvn = layers(dom, dom.HV)
    ho = hn.copy()
    if modes is None:
        modes = [(1, 1, 1.0)]
    U, V = streamfunction_transports(dom, modes, umax)
    uu, vv = layer_velocities(dom, U, V, hun, hvn)
    ww = ww_continuity(dom, uu, vv, ho, hn, dt)
    return Transport(dom, dt, uu, vv, ww, hun, hvn, ho, hn, {})


def blob_case(seed=1, imax=100, jmax=100, kmax=20) -> Transport:
    """Plain FP64 split advection of one Gaussian tracer on sigma layers with
    solid-body-like rotation (max 0.5 m/s), dx=dy=500 m, dt=30 s."""
    tr = config_advection(imax, jmax, kmax, seed, hmin=10.0, hmax=40.0,
                          coords="sigma", umax=0.5)
    ci = (imax + 1) / 2.0 + 20.0 * imax / 100.0
    cj = (jmax + 1) / 2.0
    blob = gaussian_blob(tr.dom, ci, cj, 8.0 * imax / 100.0)
    tr.tracers["c"] = blob
    return tr


def adjoint_case(seed=4, imax=256, jmax=256, kmax=30) -> Transport:
    """Reverse-mode workload: sigma layers, the Gaussian-blob flow and tracer
    of blob_case at 256x256x30, plus T and S for the internal-pressure term of
    the cost."""
    tr = blob_case(seed=seed, imax=imax, jmax=jmax, kmax=kmax)
    T, S = temperature_salinity(tr, seed)
    tr.tracers["T"] = T
    tr.tracers["S"] = S
    return tr


def seeded_modes(seed, n=4, mmax=3):
    u = uniform((n, 3), seed * 104729 + 5)
    return [(1 + int(u[q, 0] * mmax), 1 + int(u[q, 1] * mmax),
             0.25 + 0.75 * u[q, 2]) for q in range(n)]


def temperature_salinity(tr: Transport, seed):
    """T = 5 + 15 (zc+H)/H + N(0,0.2); S = 5 + 25 (i-imin)/(imax-imin);
    land and k=0 hold the -9999 sentinel (temperature.F90:282-283)."""
    dom = tr.dom
    kmax = dom.kmax
    zc = -dom.H[None] + np.cumsum(tr.hn[1:], axis=0) - 0.5 * tr.hn[1:]
    with np.errstate(divide="ignore", invalid="ignore"):
        frac = (zc + dom.H[None]) / np.where(dom.H > 0, dom.H, 1.0)[None]
    T = np.zeros(dom.shape3)
    T[1:] = 5.0 + 15.0 * frac + 0.2 * normal((kmax, dom.NY, dom.NX), seed * 31 + 3)
    ii = np.arange(dom.NX)[None, :] - HALO + dom.imin
    S = np.zeros(dom.shape3)
    S[1:] = (5.0 + 25.0 * (ii - dom.imin) / float(dom.imax - dom.imin))[None] \
        * np.ones((kmax, dom.NY, 1))
    land = dom.az != 1
    T[:, land] = -9999.0
    S[:, land] = -9999.0
    T[0] = -9999.0
    S[0] = -9999.0
    return T, S


def ts_case(seed=2, imax=400, jmax=400, kmax=40) -> Transport:
    """FP64 transport step on GVC layers: T (5-20 C, linear in depth plus
    noise) and S (5-30 psu, estuarine gradient), P2-PDM advection,
    divergence-free streamfunction flow with max 1 m/s."""
    tr = config_advection(imax, jmax, kmax, seed, hmin=10.0, hmax=50.0,
                          coords="gvc", umax=1.0, modes=seeded_modes(seed))
    T, S = temperature_salinity(tr, seed)
    tr.tracers["T"] = T
    tr.tracers["S"] = S
    return tr


def passive_blobs(dom: Domain, n, seed):
    """n passive tracers: seeded U[0,1] amplitude Gaussian blobs."""
    u = uniform((n, 4), seed * 7 + 101)
    out = []
    for q in range(n):
        ci = dom.imin + (0.2 + 0.6 * u[q, 0]) * (dom.imax - dom.imin)
        cj = dom.jmin + (0.2 + 0.6 * u[q, 1]) * (dom.jmax - dom.jmin)
        sig = 3.0 + 0.1 * u[q, 2] * min(dom.imax, dom.jmax)
        out.append(gaussian_blob(dom, ci, cj, sig, amp=u[q, 3]))
    return out


MULTITRACER_SUBDOMAINS = [(64, 64), (96, 96), (128, 128), (192, 128), (256, 192),
                 (320, 256), (384, 320), (512, 384)]


def multitracer_case(seed=3, kmax=30, sizes=None, npassive=16):
    """T, S and ``npassive`` passive tracers on uneven subdomains."""
    sizes = MULTITRACER_SUBDOMAINS if sizes is None else sizes
    subs = []
    for s, (ni, nj) in enumerate(sizes):
        tr = config_advection(ni, nj, kmax, seed * 100 + s, hmin=10.0,
                              hmax=50.0, coords="gvc", umax=1.0,
                              modes=seeded_modes(seed * 100 + s))
        T, S = temperature_salinity(tr, seed * 100 + s)
        tr.tracers["T"] = T
        tr.tracers["S"] = S
        for q, b in enumerate(passive_blobs(tr.dom, npassive, seed * 100 + s)):
            tr.tracers[f"p{q}"] = b
        subs.append(tr)
    return subs
