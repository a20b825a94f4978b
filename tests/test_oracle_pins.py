"""Pins of the oracle (CPU): each check is something the mathematics fixes,
independent of how the oracle writes the routine, chosen so that a dropped
term, a wrong sign or index, or a swapped operand fails.

The reference (getm-model/code) holds no test vectors for this path and its
Fortran cannot be built here (src/futils/exceptions.F90 needs netCDF), so the
oracle is pinned by: the published limiter functions (Sweby diagram of
Superbee, van Leer's MUSCL, Leonard's ULTIMATE-QUICKEST and QUICKEST),
textbook upwind, conservation, constancy, mirror symmetry, sub-cycling
equivalence, closed-form hydrostatic integrals, and exact rational sums.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from code_amd import inputs
from oracle import pyoracle as po

U, P2, SB, MU, PDM = po.UPSTREAM, po.P2, po.SUPERBEE, po.MUSCL, po.P2_PDM


# ---------------------------------------------------------------- limiter --
def sweby(scheme, r, c):
    """published limiter functions phi(r), zero for r <= 0"""
    if r <= 0:
        return 0.0
    if scheme == SB:      # Roe (1985), piecewise form of the Sweby diagram
        if r <= 0.5:
            return 2 * r
        if r <= 1:
            return 1.0
        if r <= 2:
            return r
        return 2.0
    if scheme == MU:      # van Leer (1979)
        return max(0.0, min(2 * r, (1 + r) / 2, 2.0))
    if scheme == PDM:     # Leonard (1991) ULTIMATE-QUICKEST; GETM guards 2r/c
        x = (1 - 2 * c) / 6   # with c + 1e-10 (advection.F90:811)
        return max(0.0, min(2 * r / (c + 1e-10), 0.5 + x + (0.5 - x) * r,
                            2 / (1 - c) if c < 1 else math.inf))
    raise ValueError


@pytest.mark.parametrize("scheme", [SB, MU, PDM])
def test_limiter_against_published_functions(scheme):
    rng = np.random.default_rng(scheme)
    for _ in range(20000):
        c = rng.uniform(0.01, 0.99)
        fuu, fu, fd = rng.uniform(-1, 1, 3)
        if rng.uniform() < 0.05:
            fuu = fu  # flat upstream
        got = po.reconstruct(scheme, c, fuu, fu, fd)
        dfd, dfu = fd - fu, fu - fuu
        if dfd * dfu > 0:
            r = dfu / dfd
            want = fu + 0.5 * sweby(scheme, r, c) * (1 - c) * dfd
        else:
            want = fu
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (scheme, c, fuu, fu, fd)


def test_p2_is_quickest():
    """P2 face value = QUICKEST (Leonard 1979):
    fu + (1-c)/2 (fd-fu) - (1-c^2)/6 (fd - 2 fu + fuu)"""
    rng = np.random.default_rng(3)
    for _ in range(20000):
        c = rng.uniform(0, 1)
        fuu, fu, fd = rng.uniform(-1, 1, 3)
        want = fu + 0.5 * (1 - c) * (fd - fu) - (1 - c * c) / 6 * (fd - 2 * fu + fuu)
        got = po.reconstruct(P2, c, fuu, fu, fd)
        assert abs(got - want) <= 1e-14 * 8


def test_limiter_special_cases():
    for s in (U, P2, SB, MU, PDM):
        assert po.reconstruct(s, 0.3, 0.7, 0.7, 0.7) == 0.7        # uniform
    for s in (U, SB, MU, PDM):
        assert po.reconstruct(s, 0.3, 0.2, 0.5, 0.1) == 0.5        # extremum: upwind
    assert po.reconstruct(U, 0.3, 0.0, 0.5, 1.0) == 0.5            # upstream
    # smooth ramp (r = 1): Superbee/MUSCL give Lax-Wendroff, fu + (1-c)/2 df
    for s in (SB, MU):
        assert abs(po.reconstruct(s, 0.4, 0.0, 1.0, 2.0) - (1 + 0.5 * 0.6)) < 1e-15


# ------------------------------------------------------------- split steps --
def closed_basin(imax=30, jmax=22, kmax=5, seed=1):
    return inputs.make_domain(imax, jmax, kmax, seed, walls=3)


def rnd(shape, seed, lo=0.0, hi=1.0):
    return inputs.uniform(shape, seed, lo, hi)


def test_upwind_textbook_uniform_flow():
    """UPSTREAM, uniform U and D: f_i <- f_i - c (f_i - f_{i-1})."""
    dom = closed_basin()
    orc = po.Oracle.from_domain(dom)
    NY, NX = dom.NY, dom.NX
    f = rnd((NY, NX), 5)
    D = np.full((NY, NX), 3.0)
    Uv = np.zeros((NY, NX))
    c = 0.37
    Uv[dom.au >= 1] = c * 3.0 * dom.dx / 30.0
    fi, Di, adv = f.copy(), D.copy(), np.zeros_like(f)
    orc.adv_split_u(30.0, f, fi, Di, adv, Uv, D, 1.0, U, -1.0)
    # interior cells whose both faces are open: textbook update
    both = (dom.au[:, 1:-1] >= 1) & (dom.au[:, :-2] >= 1)
    want = f[:, 1:-1] - c * (f[:, 1:-1] - f[:, :-2])
    got = fi[:, 1:-1]
    assert np.max(np.abs(got - want)[both]) < 1e-14
    assert np.array_equal(Di[:, 1:-1][both], D[:, 1:-1][both])  # no divergence


def mirror_x(a):
    return a[..., ::-1].copy()


def mirror_face_x(a):
    """U at face i (between i and i+1) -> -U at face NX-2-i"""
    b = np.zeros_like(a)
    b[..., :-1] = -a[..., -2::-1]
    return b


@pytest.mark.parametrize("scheme", [U, P2, SB, MU, PDM])
def test_split_u_mirror_symmetry(scheme):
    """Reflecting the basin and the flow in x reflects the result: pins the
    downwind branch (fuu=f(i+2), fd=f(i)) against the upwind one."""
    dom = closed_basin(seed=2)
    NY, NX = dom.NY, dom.NX
    orc = po.Oracle.from_domain(dom)
    azm = mirror_x(dom.az)
    orcm = po.Oracle(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, azm)
    f = rnd((NY, NX), 11)
    D = rnd((NY, NX), 12, 1.0, 2.0)
    DU = rnd((NY, NX), 13, 1.0, 2.0)
    Uv = rnd((NY, NX), 14, -1, 1) * DU * 0.8 * dom.dx / 30.0 * 2
    Uv[dom.au < 1] = 0.0
    fi, Di, adv = f.copy(), D.copy(), np.zeros_like(f)
    orc.adv_split_u(30.0, f, fi, Di, adv, Uv, DU, 0.5, scheme, -1.0)
    fm, Dm = mirror_x(f), mirror_x(D)
    DUm = np.zeros_like(DU)
    DUm[:, :-1] = DU[:, -2::-1]
    fim, Dim, advm = fm.copy(), Dm.copy(), np.zeros_like(f)
    orcm.adv_split_u(30.0, fm, fim, Dim, advm, mirror_face_x(Uv), DUm, 0.5, scheme, -1.0)
    assert np.max(np.abs(mirror_x(fim) - fi)) < 1e-13
    assert np.max(np.abs(mirror_x(Dim) - Di)) < 1e-13


@pytest.mark.parametrize("scheme", [U, P2, SB, MU, PDM])
def test_split_uv_conservation_and_constancy(scheme):
    dom = closed_basin(seed=3)
    orc = po.Oracle.from_domain(dom)
    NY, NX = dom.NY, dom.NX
    wet = dom.az == 1
    for direction in ("u", "v"):
        f = rnd((NY, NX), 21)
        D = rnd((NY, NX), 22, 1.0, 2.0)
        DV = rnd((NY, NX), 23, 1.0, 2.0)
        V = rnd((NY, NX), 24, -1, 1) * DV * 0.9 * 500.0 / 30.0
        mask = dom.au if direction == "u" else dom.av
        V[mask < 1] = 0.0
        fn = orc.adv_split_u if direction == "u" else orc.adv_split_v
        fi, Di, adv = f.copy(), D.copy(), np.zeros_like(f)
        fn(30.0, f, fi, Di, adv, V, DV, 1.0, scheme, -1.0)
        m0 = math.fsum((D * f)[wet])
        m1 = math.fsum((Di * fi)[wet])
        assert abs(m1 - m0) < 1e-12 * abs(m0)
        c = np.full((NY, NX), 4.25)
        ci, Dc, advc = c.copy(), D.copy(), np.zeros_like(c)
        fn(30.0, c, ci, Dc, advc, V, DV, 1.0, scheme, -1.0)
        assert np.max(np.abs(ci - 4.25)) < 1e-13
        assert np.array_equal(Dc, Di)   # the volume update does not see f


@pytest.mark.parametrize("scheme", [SB, MU, PDM])
def test_no_new_extrema_1d(scheme):
    """TVD schemes at CFL<=1 in uniform flow create no new extrema."""
    dom = closed_basin(seed=4)
    orc = po.Oracle.from_domain(dom)
    NY, NX = dom.NY, dom.NX
    f = rnd((NY, NX), 31)
    D = np.full((NY, NX), 2.0)
    Uv = np.zeros((NY, NX))
    Uv[dom.au >= 1] = 0.9 * 2.0 * dom.dx / 30.0
    fi, Di, adv = f.copy(), D.copy(), np.zeros_like(f)
    orc.adv_split_u(30.0, f, fi, Di, adv, Uv, D, 1.0, scheme, -1.0)
    wet = dom.az == 1
    lo = np.minimum(np.minimum(f[:, :-2], f[:, 1:-1]), f[:, 2:])
    hi = np.maximum(np.maximum(f[:, :-2], f[:, 1:-1]), f[:, 2:])
    g = fi[:, 1:-1]
    w = wet[:, 1:-1]
    assert np.all(g[w] >= lo[w] - 1e-14) and np.all(g[w] <= hi[w] + 1e-14)


def column_case(kmax=12, seed=5, wscale=1.0):
    dom = closed_basin(10, 8, kmax, seed)
    sh = dom.shape3
    f = rnd(sh, seed + 1)
    h = rnd(sh, seed + 2, 1.0, 2.0)
    ww = rnd(sh, seed + 3, -1, 1) * wscale * 0.9 / 30.0
    ww[0] = 0
    ww[-1] = 0
    return dom, f, h, ww


@pytest.mark.parametrize("scheme", [U, P2, SB, MU, PDM])
def test_split_w_column_conservation_constancy(scheme):
    dom, f, h, ww = column_case()
    orc = po.Oracle.from_domain(dom)
    wet = dom.az == 1
    fi, hi, adv = f.copy(), h.copy(), np.zeros_like(f)
    orc.adv_split_w(30.0, f, fi, hi, adv, ww, 1.0, scheme, po.H_TAG, 1)
    for j, i in zip(*np.where(wet)):
        m0 = math.fsum(h[1:, j, i] * f[1:, j, i])
        m1 = math.fsum(hi[1:, j, i] * fi[1:, j, i])
        assert abs(m1 - m0) < 1e-13 * abs(m0)
    c = np.full_like(f, -1.5)
    ci, hc, advc = c.copy(), h.copy(), np.zeros_like(f)
    orc.adv_split_w(30.0, c, ci, hc, advc, ww, 1.0, scheme, po.H_TAG, 1)
    assert np.max(np.abs(ci - (-1.5))) < 1e-13


def test_split_w_upwind_textbook():
    """uniform h, upward w > 0 at every interface: f_k <- f_k - c (f_k - f_{k-1})"""
    dom, f, h, ww = column_case(kmax=9)
    orc = po.Oracle.from_domain(dom)
    h[:] = 2.0
    ww[1:-1] = 0.3 * 2.0 / 30.0
    fi, hi, adv = f.copy(), h.copy(), np.zeros_like(f)
    orc.adv_split_w(30.0, f, fi, hi, adv, ww, 1.0, U, po.H_TAG, 1)
    c = 0.3
    wet = dom.az == 1
    k = np.arange(2, dom.kmax)           # interior levels
    want = f[k] - c * (f[k] - f[k - 1])
    assert np.max(np.abs(fi[k] - want)[:, wet]) < 1e-14
    # the bottom layer loses water and tracer through its top face together:
    # its concentration is unchanged while its height drops by dt*w
    assert np.max(np.abs(fi[1] - f[1])[wet]) < 1e-14
    assert np.max(np.abs(hi[1] - (2.0 - 30.0 * ww[1]))[wet]) < 1e-14


@pytest.mark.parametrize("scheme", [U, SB, PDM])
def test_split_w_subcycling_equals_half_steps(scheme):
    """with max vertical CFL in (1, 2], itersmax>1 gives exactly two half
    steps (the CFL passed to the limiter is scaled by 1/2 either way)"""
    dom, f, h, ww = column_case(kmax=10, seed=8, wscale=1.0)
    h[:] = 1.0
    # one-signed w per column with vertical CFL 1.2..1.4: two iterations
    sgn = np.where(inputs.uniform(ww.shape[1:], 9) < 0.5, -1.0, 1.0)
    ww = sgn[None] * (1.2 + 0.2 * inputs.uniform(ww.shape, 10)) / 30.0
    ww[0] = 0
    ww[-1] = 0
    orc = po.Oracle.from_domain(dom)
    fi, hi, adv = f.copy(), h.copy(), np.zeros_like(f)
    nw = orc.adv_split_w(30.0, f, fi, hi, adv, ww, 1.0, scheme, po.H_TAG, 5)
    assert nw == 0
    g, hg, ag = f.copy(), h.copy(), np.zeros_like(f)
    orc.adv_split_w(30.0, g, g, hg, ag, ww, 0.5, scheme, po.H_TAG, 1)
    orc.adv_split_w(30.0, g, g, hg, ag, ww, 0.5, scheme, po.H_TAG, 1)
    wet = dom.az == 1
    # identical except where the second sweep's CFL (from the updated h)
    # would have asked for a third iteration (restart); check it did not
    assert np.array_equal(fi[:, wet], g[:, wet])
    assert np.array_equal(hi[:, wet], hg[:, wet])


def test_upstream_2dh_bounds_and_textbook():
    dom = closed_basin(seed=9)
    orc = po.Oracle.from_domain(dom)
    NY, NX = dom.NY, dom.NX
    f = rnd((NY, NX), 41)
    D = rnd((NY, NX), 42, 1, 2)
    Uv = rnd((NY, NX), 43, -1, 1) * 0.4 * 500 / 30
    Vv = rnd((NY, NX), 44, -1, 1) * 0.4 * 500 / 30
    Uv[dom.au < 1] = 0
    Vv[dom.av < 1] = 0
    fi, Di, adv = f.copy(), D.copy(), np.zeros_like(f)
    orc.adv_upstream_2dh(30.0, f, fi, Di, adv, Uv, Vv, D, D, D, -1.0)
    wet = dom.az == 1
    for j, i in zip(*np.where(wet)):
        nb = [f[j, i + q] for q in (-1, 0, 1) if dom.az[j, i + q] == 1] + \
             [f[j + q, i] for q in (-1, 1) if dom.az[j + q, i] == 1]
        assert min(nb) - 1e-15 <= fi[j, i] <= max(nb) + 1e-15
    # the flux-form update before the clip, from the textbook upwind fluxes
    uf = np.where(Uv > 0, Uv * f, Uv * np.roll(f, -1, axis=1))
    vf = np.where(Vv > 0, Vv * f, Vv * np.roll(f, -1, axis=0))
    div = ((uf - np.roll(uf, 1, axis=1)) * dom.dy + (vf - np.roll(vf, 1, axis=0)) * dom.dx) \
        / (dom.dx * dom.dy)
    assert np.max(np.abs(adv - div)[wet]) < 1e-12


def test_nosplit_upstream_2dh_discards_the_clip():
    """3-D NOSPLIT with UPSTREAM_2DH rebuilds the tracer from the advective
    term, f = (ho f - dt adv)/hi (advection_3d.F90:338-350,416-423), so the
    clip of adv_upstream_2dh does not survive: on water cells the result is
    the textbook upwind flux form, including cells where the clip would
    have changed the value (ww = 0, vertical UPSTREAM is then the identity).
    The upwind flux form is a convex combination of a cell and its upwind
    neighbours while the outflow stays below the cell's volume, so the
    transports here reach Courant numbers above 1 for the clip to act."""
    dom = closed_basin(imax=24, jmax=20, kmax=3, seed=9)
    orc = po.Oracle.from_domain(dom)
    sh = dom.shape3
    dt = 30.0
    f = rnd(sh, 51)
    ho = rnd(sh, 52, 1.0, 2.0)
    hu = rnd(sh, 53, 1.0, 2.0)
    hv = rnd(sh, 54, 1.0, 2.0)
    uu = rnd(sh, 55, -1, 1) * 1.5 * dom.dx / dt
    vv = rnd(sh, 56, -1, 1) * 1.5 * dom.dy / dt
    uu[:, dom.au < 1] = 0.0
    vv[:, dom.av < 1] = 0.0
    ww = np.zeros(sh)
    g, hires = f.copy(), np.zeros(sh)
    rc, _ = orc.do_advection_3d(dt, g, uu, vv, ww, hu, hv, ho, ho, po.NOSPLIT, po.UPSTREAM_2DH,
                                U, -1.0, po.H_TAG, 1, hires, None)
    assert rc == 0
    wet = dom.az == 1
    clipped = 0
    for k in range(1, dom.kmax + 1):
        fk, U_, V_ = f[k], uu[k], vv[k]
        uf = np.where(U_ > 0, U_ * fk, U_ * np.roll(fk, -1, axis=1))
        vf = np.where(V_ > 0, V_ * fk, V_ * np.roll(fk, -1, axis=0))
        div = ((uf - np.roll(uf, 1, axis=1)) * dom.dy + (vf - np.roll(vf, 1, axis=0)) * dom.dx) \
            / (dom.dx * dom.dy)
        dU = ((U_ - np.roll(U_, 1, axis=1)) * dom.dy + (V_ - np.roll(V_, 1, axis=0)) * dom.dx) \
            / (dom.dx * dom.dy)
        hi = ho[k] - dt * dU
        want = (ho[k] * fk - dt * div) / hi
        assert np.max((np.abs(g[k] - want) / np.maximum(1.0, np.abs(want)))[wet]) < 1e-12, k
        # the 2-D routine alone clips: its result differs where the flux
        # form leaves the neighbours' range
        fi, Di, adv = fk.copy(), ho[k].copy(), np.zeros_like(fk)
        orc.adv_upstream_2dh(dt, fk, fi, Di, adv, U_, V_, ho[k], hu[k], hv[k], -1.0)
        clipped += int((np.abs(fi - want) > 1e-9 * np.maximum(1.0, np.abs(want)))[wet].sum())
    assert clipped > 0


# ---------------------------------------------------------- composition --
def test_halfsplit_is_the_five_substeps_and_conserves():
    tr = inputs.blob_case(imax=40, jmax=36, kmax=6)
    d = tr.dom
    orc = po.Oracle.from_domain(d)
    f = tr.tracers["c"].copy()
    hires = np.zeros(d.shape3)
    rc, _ = orc.do_advection_3d(tr.dt, f, tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn,
                                po.HALFSPLIT, SB, SB, -1.0, po.H_TAG, 1, hires, None)
    assert rc == 0
    g = tr.tracers["c"].copy()
    hi = tr.ho.copy()
    adv = np.zeros(d.shape3)
    for k in range(1, d.kmax + 1):
        orc.adv_split_u(tr.dt, g[k], g[k], hi[k], adv[k], tr.uu[k], tr.hun[k], 0.5, SB, -1.0)
    for k in range(1, d.kmax + 1):
        orc.adv_split_v(tr.dt, g[k], g[k], hi[k], adv[k], tr.vv[k], tr.hvn[k], 0.5, SB, -1.0)
    orc.adv_split_w(tr.dt, g, g, hi, adv, tr.ww, 1.0, SB, po.H_TAG, 1)
    for k in range(1, d.kmax + 1):
        orc.adv_split_v(tr.dt, g[k], g[k], hi[k], adv[k], tr.vv[k], tr.hvn[k], 0.5, SB, -1.0)
    for k in range(1, d.kmax + 1):
        orc.adv_split_u(tr.dt, g[k], g[k], hi[k], adv[k], tr.uu[k], tr.hun[k], 0.5, SB, -1.0)
    assert np.array_equal(f, g) and np.array_equal(hires, hi)
    wet = d.az == 1
    # non-divergent depth-integrated transport: the volume trajectory ends on hn
    assert np.max(np.abs(hires - tr.hn)[1:, wet]) < 1e-12
    m0 = math.fsum((tr.hn * tr.tracers["c"])[1:, wet].ravel())
    m1 = math.fsum((hires * f)[1:, wet].ravel())
    assert abs(m1 - m0) < 1e-13 * m0


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("scheme", [U, SB, PDM])
def test_halfsplit_open_domain_mirror_symmetry(scheme, axis):
    """Water up to and into the halo: the reference's horizontal flux and
    update ranges ([isub, NX-2-isub] for faces, [1+isub, NX-2-isub] for
    cells, adv_split_u.F90:205-286, adv_split_v.F90:63-136) are symmetric
    under a reflection, so a HALFSPLIT step commutes with reflecting the
    basin, the tracer and the transports in x (y); a bound that is off by
    one row updates one edge and not the other.  ww = 0: the vertical
    sub-step's column range stops at imax+1 on one side only
    (adv_split_w.F90:85-305), by the reference's design."""
    imax, jmax, kmax = 22, 17, 4
    NX, NY = imax + 4, jmax + 4
    az = np.ones((NY, NX), np.int32)
    isl = inputs.uniform((NY, NX), 91) < 0.1
    isl[:3, :] = isl[-3:, :] = False
    isl[:, :3] = isl[:, -3:] = False
    az[isl] = 0
    au, av, ax = inputs.masks_from_az(az)
    H = inputs.uniform((NY, NX), 92, 10.0, 40.0)
    HU, HV = H.copy(), H.copy()
    HU[:, :-1] = 0.5 * (H[:, :-1] + H[:, 1:])
    HV[:-1, :] = 0.5 * (H[:-1, :] + H[1:, :])
    dom = inputs.Domain(imax, jmax, kmax, 500.0, 400.0, H, HU, HV, az, au, av, ax)
    sh = dom.shape3
    dt = 30.0
    f = rnd(sh, 93)
    ho = rnd(sh, 94, 0.5, 2.0)
    hu = rnd(sh, 95, 0.5, 2.0)
    hv = rnd(sh, 96, 0.5, 2.0)
    uu = rnd(sh, 97, -1, 1) * hu * 0.8 * dom.dx / dt
    vv = rnd(sh, 98, -1, 1) * hv * 0.8 * dom.dy / dt
    ww = np.zeros(sh)
    ax_ = -1 if axis == "x" else -2
    n = NX if axis == "x" else NY

    def m(a):
        return np.flip(a, axis=ax_).copy()

    def mface(a, sign):  # face i -> face n-2-i
        b = np.zeros_like(a)
        idx = [slice(None)] * a.ndim
        idx[ax_] = slice(0, n - 1)
        b[tuple(idx)] = sign * np.take(a, np.arange(n - 2, -1, -1), axis=ax_)
        return b

    def step(o, f_, uu_, vv_, hu_, hv_, ho_):
        g, hires = f_.copy(), np.zeros(sh)
        rc, _ = o.do_advection_3d(dt, g, uu_, vv_, ww, hu_, hv_, ho_, ho_, po.HALFSPLIT, scheme,
                                  scheme, -1.0, po.H_TAG, 1, hires, None)
        assert rc == 0
        return g, hires

    g, hires = step(po.Oracle.from_domain(dom), f, uu, vv, hu, hv, ho)
    if axis == "x":
        args = (mface(uu, -1.0), m(vv), mface(hu, 1.0), m(hv))
    else:
        args = (m(uu), mface(vv, -1.0), m(hu), mface(hv, 1.0))
    gm, hm = step(po.Oracle(imax, jmax, kmax, dom.dx, dom.dy, m(az)), m(f), *args[:2], *args[2:],
                  m(ho))
    assert np.max(np.abs(m(gm) - g)) < 1e-13 and np.max(np.abs(m(hm) - hires)) < 1e-13
    # the step reached the halo rows and columns
    assert np.max(np.abs(g - f)[1:, :2, :]) > 0 and np.max(np.abs(g - f)[1:, :, :2]) > 0


def test_error_codes():
    dom = closed_basin()
    orc = po.Oracle.from_domain(dom)
    z = np.zeros(dom.shape3)
    args = [z.copy() for _ in range(8)]
    assert orc.do_advection_3d(30.0, *args, 2, 4, 4, -1.0, 99)[0] == 1           # tag
    assert orc.do_advection_3d(30.0, *args, 7, 4, 4, -1.0, 10)[0] == 2           # split
    assert orc.do_advection_3d(30.0, *args, 2, 2, 4, -1.0, 10)[0] == 4           # 2DH+split
    assert orc.do_advection_3d(30.0, *args, 2, 12, 4, -1.0, 10)[0] == 3          # scheme


# --------------------------------------------------- EOS and pressure -----
def flat_basin(kmax=6):
    dom = inputs.make_domain(20, 16, kmax, 3, hmin=25.0, hmax=25.0, walls=2)
    hn = inputs.sigma_layers(dom, dom.H)
    hun = inputs.sigma_layers(dom, dom.HU)
    hvn = inputs.sigma_layers(dom, dom.HV)
    return dom, hn, hun, hvn


def test_eos_linear_properties():
    dom, hn, _, _ = flat_basin()
    orc = po.Oracle.from_domain(dom)
    sh = dom.shape3
    T0, S0, a, b = inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0
    T = np.full(sh, T0)
    S = np.full(sh, S0)
    rho, buoy = np.zeros(sh), np.zeros(sh)
    orc.eos_linear(T, S, rho, buoy, T0, S0, a, b)
    assert np.all(buoy[1:] == 0.0) and np.all(rho[1:] == 1025.0)
    T2 = T + 1.0
    orc.eos_linear(T2, S, rho, buoy, T0, S0, a, b)
    g = float(np.float32(9.81))
    assert np.allclose(buoy[1:], -g / 1025.0 * a, rtol=1e-12, atol=0)
    assert np.all(buoy[0] == 0.0)  # level 0 is not touched


@pytest.mark.parametrize("method", ["bm", "sm"])
def test_ip_constant_buoyancy_vanishes(method):
    dom = inputs.make_domain(24, 20, 7, 5, hmin=10, hmax=40, walls=3)
    orc = po.Oracle.from_domain(dom)
    sse = inputs.uniform((dom.NY, dom.NX), 77, -0.3, 0.3)
    hn = inputs.gvc_layers(dom, dom.H + sse)
    hun = inputs.gvc_layers(dom, dom.HU)
    hvn = inputs.gvc_layers(dom, dom.HV)
    buoy = np.full(dom.shape3, -0.0123)
    zz, ix, iy = (np.zeros(dom.shape3) for _ in range(3))
    if method == "bm":
        sse[:] = 0.0
        hn = inputs.gvc_layers(dom, dom.H)
        orc.ip_blumberg_mellor(dom.H, hn, hun, hvn, buoy, zz, ix, iy)
    else:
        orc.ip_shchepetkin_mcwilliams(dom.H, hn, hun, hvn, buoy, sse, zz, ix, iy)
    assert np.max(np.abs(ix)) < 1e-14 and np.max(np.abs(iy)) < 1e-14


@pytest.mark.parametrize("method", ["bm", "sm"])
def test_ip_linear_buoyancy_hydrostatic_integral(method):
    """flat bottom, uniform layers, b = a_x x + a_y y:
    idpdx_k = h_k * a_x * (depth of the centre of layer k)"""
    dom, hn, hun, hvn = flat_basin(kmax=8)
    orc = po.Oracle.from_domain(dom)
    ax_, ay_ = 3e-7, -2e-7
    NY, NX = dom.NY, dom.NX
    xx = np.arange(NX)[None, :] * dom.dx
    yy = np.arange(NY)[:, None] * dom.dy
    buoy = np.zeros(dom.shape3)
    buoy[1:] = (ax_ * xx + ay_ * yy)[None]
    zz, ix, iy = (np.zeros(dom.shape3) for _ in range(3))
    if method == "bm":
        orc.ip_blumberg_mellor(dom.H, hn, hun, hvn, buoy, zz, ix, iy)
    else:
        orc.ip_shchepetkin_mcwilliams(dom.H, hn, hun, hvn, buoy, np.zeros((NY, NX)), zz, ix, iy)
    hk = 25.0 / dom.kmax
    depth = (dom.kmax - np.arange(1, dom.kmax + 1) + 0.5) * hk
    inner = (dom.au >= 1) & (dom.av >= 1)
    inner[:, :] &= np.ones_like(inner)
    for k in range(1, dom.kmax + 1):
        wx = hk * ax_ * depth[k - 1]
        wy = hk * ay_ * depth[k - 1]
        sel = inner.copy()
        sel[:2, :] = sel[-2:, :] = False
        sel[:, :2] = sel[:, -2:] = False
        assert np.max(np.abs(ix[k][sel] - wx)) < 1e-12 * abs(wx) + 1e-20, (method, k)
        assert np.max(np.abs(iy[k][sel] - wy)) < 1e-12 * abs(wy) + 1e-20, (method, k)


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("method", ["bm", "sm"])
def test_ip_mirror_symmetry_with_coastlines(method, axis):
    """Reflecting the basin, its islands and every field in x (y) reflects
    the pressure gradients: idpdx (idpdy) changes sign at the mirrored U (V)
    point NX-2-i (NY-2-j), the other gradient is carried over.  Islands and
    land sentinels in the buoyancy make this pin the coastline stencils'
    indices and signs, which no closed form covers (Shchepetkin-McWilliams
    reads land values there: ~200 of its gradients change with the
    sentinel); both runs do the same operations on negated or swapped
    operands, so the bound is 1e-13 of the largest gradient."""
    imax, jmax, kmax = 26, 18, 6
    dom0 = inputs.make_domain(imax, jmax, kmax, 7, hmin=10.0, hmax=40.0, walls=2)
    NY, NX = dom0.NY, dom0.NX
    az = dom0.az.copy()
    isl = inputs.uniform((NY, NX), 71) < 0.12
    isl[:, :4] = isl[:, -4:] = False
    isl[:4, :] = isl[-4:, :] = False
    az[isl & (az == 1)] = 0
    au, av, ax = inputs.masks_from_az(az)
    H = dom0.H.copy()
    H[az == 0] = -10.0
    HU, HV = H.copy(), H.copy()
    HU[:, :-1] = 0.5 * (H[:, :-1] + H[:, 1:])
    HV[:-1, :] = 0.5 * (H[:-1, :] + H[1:, :])
    dom = inputs.Domain(imax, jmax, kmax, dom0.dx, dom0.dy, H, HU, HV, az, au, av, ax)
    sse = inputs.uniform((NY, NX), 72, -0.2, 0.2) if method == "sm" else np.zeros((NY, NX))
    hn = inputs.gvc_layers(dom, np.where(az > 0, H, 10.0) + sse)
    hun = inputs.gvc_layers(dom, np.where(au > 0, HU, 10.0))
    hvn = inputs.gvc_layers(dom, np.where(av > 0, HV, 10.0))
    buoy = inputs.uniform(dom.shape3, 73, -0.02, 0.01)
    buoy[:, az == 0] = 3.7   # the land sentinel's buoyancy
    buoy[0] = 3.7

    def run(o, H_, hn_, hun_, hvn_, b_, s_):
        zz_, ix_, iy_ = (np.zeros(dom.shape3) for _ in range(3))
        if method == "bm":
            o.ip_blumberg_mellor(H_, hn_, hun_, hvn_, b_, zz_, ix_, iy_)
        else:
            o.ip_shchepetkin_mcwilliams(H_, hn_, hun_, hvn_, b_, s_, zz_, ix_, iy_)
        return ix_, iy_

    ix, iy = run(po.Oracle.from_domain(dom), H, hn, hun, hvn, buoy, sse)
    ax_ = -1 if axis == "x" else -2
    n = NX if axis == "x" else NY

    def m(a):  # a cell field
        return np.flip(a, axis=ax_).copy()

    def mface(a):  # a face field of the mirrored direction: face i -> face n-2-i
        b = np.zeros_like(a)
        src = np.take(a, np.arange(n - 2, -1, -1), axis=ax_)
        idx = [slice(None)] * a.ndim
        idx[ax_] = slice(0, n - 1)
        b[tuple(idx)] = src
        return b
    hun_m, hvn_m = (mface(hun), m(hvn)) if axis == "x" else (m(hun), mface(hvn))
    ixm, iym = run(po.Oracle(imax, jmax, kmax, dom.dx, dom.dy, m(az)), m(H), m(hn), hun_m,
                   hvn_m, m(buoy), m(sse))
    scale = max(np.abs(ix).max(), np.abs(iy).max())
    assert scale > 0
    # faces 2..n-4 of the mirrored direction are computed in both runs (the
    # loops end at imax+HALO-2)
    flip, keep = (ix, iy) if axis == "x" else (iy, ix)
    flipm, keepm = (ixm, iym) if axis == "x" else (iym, ixm)
    sel = [slice(1, None), slice(None), slice(None)]
    selm = list(sel)
    sel[ax_] = slice(2, n - 3)
    selm[ax_] = slice(n - 4, 1, -1)
    e_flip = np.abs(flipm[tuple(selm)] + flip[tuple(sel)]).max()
    e_keep = np.abs(m(keepm)[1:] - keep[1:]).max()
    assert e_flip <= 1e-13 * scale and e_keep <= 1e-13 * scale, (method, axis, e_flip, e_keep)


# ------------------------------------------------------------ halo & budget --
def test_tile_exchange_decomposed_equals_global():
    """cut a global field into tiles with decomp_1d, blank the ghost cells,
    exchange: every ghost cell with a neighbour gets the global value"""
    ni_g, nj_g, nk, px, py = 23, 17, 3, 3, 2
    g = inputs.uniform((nk, nj_g + 4, ni_g + 4), 5)
    xs = [po.decomp_1d(ni_g, px, m) for m in range(px)]
    ys = [po.decomp_1d(nj_g, py, m) for m in range(py)]
    assert xs[0][0] == 1 and xs[-1][1] == ni_g
    assert all(xs[m][1] + 1 == xs[m + 1][0] for m in range(px - 1))
    assert max(e - s for s, e in xs) - min(e - s for s, e in xs) <= 1
    tiles, ni, nj, want = [], [], [], []
    for ty in range(py):
        for tx in range(px):
            (i0, i1), (j0, j1) = xs[tx], ys[ty]
            ref = g[:, j0 - 1:j1 + 4, i0 - 1:i1 + 4].copy()   # with 2-cell ghosts
            t = ref.copy()
            t[:, :2, :] = 0
            t[:, -2:, :] = 0
            t[:, :, :2] = 0
            t[:, :, -2:] = 0
            tiles.append(t)
            want.append(ref)
            ni.append(i1 - i0 + 1)
            nj.append(j1 - j0 + 1)
    po.exchange_tiles(px, py, ni, nj, nk, tiles)
    for q, (t, ref) in enumerate(zip(tiles, want)):
        tx, ty = q % px, q // px
        sel = np.zeros(t.shape[1:], bool)
        sel[2:-2, 2:-2] = True
        if tx > 0:
            sel[2:-2, :2] = True
        if tx < px - 1:
            sel[2:-2, -2:] = True
        if ty > 0:
            sel[:2, 2:-2] = True
        if ty < py - 1:
            sel[-2:, 2:-2] = True
        if tx > 0 and ty > 0:
            sel[:2, :2] = True
        if tx < px - 1 and ty > 0:
            sel[:2, -2:] = True
        if tx > 0 and ty < py - 1:
            sel[-2:, :2] = True
        if tx < px - 1 and ty < py - 1:
            sel[-2:, -2:] = True
        assert np.array_equal(t[:, sel], ref[:, sel]), q
        assert np.all(t[:, ~sel] == 0)


def test_mirror_halo_3d():
    dom = closed_basin()
    orc = po.Oracle.from_domain(dom)
    f = inputs.uniform(dom.shape3, 3)
    g = f.copy()
    orc.update_3d_halo(g, g, 1)
    w = f.copy()
    w[:, :, 1] = w[:, :, 2]
    w[:, :, -2] = w[:, :, -3]
    w[:, 1, :] = w[:, 2, :]
    w[:, -2, :] = w[:, -3, :]
    assert np.array_equal(g, w)
    h = f.copy()
    orc.update_3d_halo(h, h, 0)
    assert np.array_equal(h, f)


def test_budget_exact_sum():
    dom = closed_basin(9, 7, 3, 2)
    orc = po.Oracle.from_domain(dom)
    h = inputs.uniform(dom.shape3, 1, 0.5, 3.0)
    f = inputs.uniform(dom.shape3, 2, -1e3, 1e3)
    m, v = orc.budget(h, f)
    wet = dom.az == 1
    exact = sum(Fraction(a) * Fraction(b) for a, b in zip(h[1:, wet].ravel(), f[1:, wet].ravel()))
    exact *= Fraction(dom.dx) * Fraction(dom.dy)
    assert abs(Fraction(m) - exact) <= abs(exact) * Fraction(1, 2 ** 52)
    vol = sum(Fraction(a) for a in h[1:, wet].ravel()) * Fraction(dom.dx * dom.dy)
    assert abs(Fraction(v) - vol) <= vol * Fraction(1, 2 ** 52)
