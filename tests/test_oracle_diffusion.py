"""Pins of the oracle's vertical diffusion (CPU): getm_tridiagonal,
do_salinity's and do_temperature's implicit diffusion with the short-wave
source.

* tridiagonal: the dense solution of the same system (numpy.linalg.solve)
  for every size, including the smallest (two rows) and a sub-range fi..lt.
* diffusion: the Crank-Nicolson matrix conserves the column content
  sum(hn*f) (its columns sum to hn); a uniform profile stays uniform; the
  fully implicit scheme (cnpar = 1) creates no new extrema; a two-layer
  column equals the closed-form 2x2 solution; land, halo columns and
  kmax = 1 are left alone.
* temperature: the short-wave profile is the two-band exponential of the
  depth below the surface (Paulson and Simpson 1977); the heat content
  changes by exactly dt * (surface flux + absorbed short-wave) (with the
  bottom reflection when enabled); the primitive sea-ice rule drops a
  cooling surface flux.
"""
import numpy as np
import pytest

from code_amd import inputs
from oracle import pyoracle as po

AVMOLS = float(np.float32(1.1e-9))
AVMOLT = float(np.float32(1.4e-7))
RHO_0_CPI = 1.0 / (1025.0 * 3985.0)


def dense(kmax, fi, lt, au, bu, cu, du):
    n = lt - fi + 1
    M = np.zeros((n, n))
    for r, i in enumerate(range(fi, lt + 1)):
        M[r, r] = bu[i]
        if i > fi:
            M[r, r - 1] = au[i]
        if i < lt:
            M[r, r + 1] = cu[i]
    return np.linalg.solve(M, du[fi:lt + 1])


@pytest.mark.parametrize("kmax,fi,lt", [(2, 1, 2), (3, 1, 3), (10, 1, 10), (40, 1, 40),
                                        (12, 0, 12), (12, 3, 9)])
def test_tridiagonal_matches_dense_solve(kmax, fi, lt):
    rng = np.random.default_rng(kmax * 100 + fi)
    for _ in range(20):
        au = rng.uniform(-1, 0, kmax + 1)
        cu = rng.uniform(-1, 0, kmax + 1)
        bu = 2.5 + rng.uniform(0, 1, kmax + 1)
        du = rng.uniform(-3, 3, kmax + 1)
        got = po.tridiagonal(kmax, fi, lt, au, bu, cu, du)
        want = dense(kmax, fi, lt, au, bu, cu, du)
        assert np.allclose(got[fi:lt + 1], want, rtol=1e-13, atol=1e-14)


def column_case(kmax=8, seed=5):
    dom = inputs.make_domain(9, 7, kmax, seed, walls=1)
    orc = po.Oracle.from_domain(dom)
    sh = dom.shape3
    hn = inputs.uniform(sh, seed + 1, 0.3, 3.0)
    nuh = inputs.uniform(sh, seed + 2, 1e-5, 5e-2)
    f = inputs.uniform(sh, seed + 3, 5.0, 30.0)
    return dom, orc, hn, nuh, f


def water(dom):
    w = np.zeros(dom.az.shape, bool)
    w[2:-2, 2:-2] = dom.az[2:-2, 2:-2] == 1
    return w


@pytest.mark.parametrize("cnpar", [0.5, 0.9, 1.0])
def test_salinity_diffusion_conserves_and_bounds(cnpar):
    dom, orc, hn, nuh, S = column_case()
    S0 = S.copy()
    orc.salinity_diffusion(600.0, cnpar, hn, nuh, S)
    w = water(dom)
    m0 = (hn[1:] * S0[1:]).sum(0)
    m1 = (hn[1:] * S[1:]).sum(0)
    assert np.allclose(m1[w], m0[w], rtol=1e-13, atol=0)
    assert np.array_equal(S[:, ~w], S0[:, ~w]) and np.array_equal(S[0], S0[0])
    assert not np.allclose(S[1:, w], S0[1:, w])  # it did diffuse
    if cnpar == 1.0:  # implicit Euler: an M-matrix, no new extrema
        assert np.all(S[1:, w] <= S0[1:, w].max(0) + 1e-12)
        assert np.all(S[1:, w] >= S0[1:, w].min(0) - 1e-12)
    U = np.full_like(S0, 17.0)
    orc.salinity_diffusion(600.0, cnpar, hn, nuh, U)
    assert np.allclose(U, 17.0, rtol=1e-14, atol=0)


@pytest.mark.parametrize("cnpar", [0.5, 1.0])
def test_two_layer_closed_form(cnpar):
    dom, orc, hn, nuh, S = column_case(kmax=2)
    S0 = S.copy()
    dt = 900.0
    orc.salinity_diffusion(dt, cnpar, hn, nuh, S)
    w = water(dom)
    for j, i in zip(*np.nonzero(w)):
        h1, h2 = hn[1, j, i], hn[2, j, i]
        k = 2 * dt * (nuh[1, j, i] + AVMOLS) / (h1 + h2)
        an, ao = cnpar * k, (1 - cnpar) * k
        M = np.array([[h1 + an, -an], [-an, h2 + an]])
        rhs = np.array([S0[1, j, i] * (h1 - ao) + S0[2, j, i] * ao,
                        S0[2, j, i] * (h2 - ao) + S0[1, j, i] * ao])
        assert np.allclose(S[1:, j, i], np.linalg.solve(M, rhs), rtol=1e-13, atol=0)


def test_kmax_one_is_untouched():
    dom, orc, hn, nuh, S = column_case(kmax=1)
    S0 = S.copy()
    orc.salinity_diffusion(600.0, 0.5, hn, nuh, S)
    assert np.array_equal(S, S0)


def temperature_inputs(dom, seed=11, swr_max=400.0):
    sh2 = dom.az.shape
    swr = inputs.uniform(sh2, seed, 0.0, swr_max)
    shf = inputs.uniform(sh2, seed + 1, -200.0, 100.0)
    A = inputs.uniform(sh2, seed + 2, 0.5, 0.7)
    g1 = inputs.uniform(sh2, seed + 3, 0.3, 1.5)
    g2 = inputs.uniform(sh2, seed + 4, 5.0, 20.0)
    dry_z = np.ones(sh2)
    return swr, shf, A, g1, g2, dry_z


@pytest.mark.parametrize("refl", [-1.0, 0.3])
def test_temperature_radiation_and_heat_budget(refl):
    dom, orc, hn, nuh, T = column_case(kmax=7, seed=21)
    hn = hn * 0.5  # shallow enough for the bottom reflection to apply
    S = inputs.uniform(dom.shape3, 3, 20.0, 35.0)
    swr, shf, A, g1, g2, dry_z = temperature_inputs(dom)
    rad = np.zeros(dom.shape3)
    T0 = T.copy()
    dt, kmax = 600.0, dom.kmax
    orc.temperature_diffusion(dt, 0.6, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T,
                              swr_bot_refl_frac=refl, swr_min_bot_frac=0.01)
    w = water(dom)
    # the short-wave profile: depth of interface k below the surface
    depth = np.zeros(dom.shape3)
    depth[:kmax] = np.cumsum(hn[:0:-1], axis=0)[::-1]
    want = swr[None] * (A[None] * np.exp(-depth / g1[None])
                        + (1 - A[None]) * np.exp(-depth / g2[None]))
    assert np.allclose(rad[:, w], want[:, w], rtol=1e-14, atol=1e-12)
    # heat content: dt * (absorbed short-wave + surface flux), per unit rho0 cp
    top, bot = rad[kmax].copy(), rad[0].copy()
    if refl > 0:
        with np.errstate(invalid="ignore", divide="ignore"):
            on = (bot / top > float(np.float32(0.01)))
        assert on[w].any()
        D = hn[1:].sum(0)
        r = bot * refl
        top = np.where(on, top - r * (A * np.exp(-D / g1) + (1 - A) * np.exp(-D / g2)), top)
        bot = np.where(on, bot - r, bot)
    shf_eff = np.where(T0[kmax] <= -0.0575 * S[kmax], np.maximum(shf, 0.0), shf)
    gain = dt * (top + shf_eff - bot) * RHO_0_CPI
    m0 = (hn[1:] * T0[1:]).sum(0)
    m1 = (hn[1:] * T[1:]).sum(0)
    assert np.allclose(m1[w] - m0[w], gain[w], rtol=1e-9, atol=1e-12)


def test_sea_ice_rule_drops_cooling():
    dom, orc, hn, nuh, T = column_case(kmax=5, seed=31)
    S = np.full(dom.shape3, 35.0)
    T[:] = -2.1  # below -0.0575 * 35 = -2.0125: freezing
    swr, shf, A, g1, g2, dry_z = temperature_inputs(dom, swr_max=0.0)
    shf[:] = -300.0
    rad = np.zeros(dom.shape3)
    T0 = T.copy()
    orc.temperature_diffusion(600.0, 0.5, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T)
    w = water(dom)
    assert np.allclose(T[1:, w], T0[1:, w], rtol=1e-14, atol=0)  # no flux, no radiation
    T2 = np.full(dom.shape3, 5.0)
    orc.temperature_diffusion(600.0, 0.5, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T2)
    assert np.all(T2[dom.kmax, w] < 5.0)  # above freezing the cooling acts
