"""Pins of the oracle's full equation of state (CPU).

* Jackett et al. (2006): the check values the reference quotes in
  src/3d/eqstate.F90:356 and :448-452 (rho, d rho/dS, d rho/d theta at
  S = 20, theta = 20, p = 1000 dbar).  The value of :343-347
  (S = 35, T = 25, p = 10000 -> 1062.53817) is EOS-80's in-situ check value;
  this function takes potential temperature, so it does not reproduce it and
  it is not used.
* UNESCO 1980: the published one-atmosphere check values (Millero and
  Poisson 1981: 1027.67547 at S = 35, T = 5; 1023.34306 at S = 35, T = 25),
  the density maximum of pure water (999.975 kg/m3 at 3.98 C), and
  agreement with Jackett et al. over the oceanic range (the two fits differ
  by < 0.02 kg/m3).
* eosall_from_theta: its derivatives equal central differences of
  rho_from_theta at several pressures (pins every derivative term against the
  density polynomial, itself pinned by the check value).
* do_eqstate: the column logic (pressure accumulated from the surface, the
  interface means, land columns, kmax, the linear fill) against those pinned
  point functions on hand-built columns.
"""
import json
from pathlib import Path

import numpy as np
import pytest

from code_amd import inputs
from oracle import pyoracle as po


GOLDEN = json.loads((Path(__file__).parent / "golden" / "eqstate_check_values.json").read_text())


def test_jackett_reference_check_values():
    for g in GOLDEN["rho_from_theta"]:
        dens0, densp = po.rho_from_theta(g["s"], g["th"], g["p"])
        assert abs(densp - g["densp"]) <= g["tol_rel"] * g["densp"], g["cite"]
        assert dens0 < densp  # compression raises the density
    for g in GOLDEN["eosall_from_theta"]:
        rho_s, rho_th = po.eosall_from_theta(g["s"], g["th"], g["p"])
        assert abs(rho_s - g["rho_s"]) <= g["tol_abs"], g["cite"]
        assert abs(rho_th - g["rho_th"]) <= g["tol_abs"], g["cite"]


def test_unesco80_published_values():
    for g in GOLDEN["rho_from_theta_unesco80"]:
        assert abs(po.rho_unesco80(g["T"], g["S"]) - g["rho"]) <= g["tol_abs"], g
    T = np.linspace(3.9, 4.1, 2001)
    r = np.array([po.rho_unesco80(t, 0.0) for t in T])
    assert abs(T[r.argmax()] - 3.98) <= 0.01
    assert abs(r.max() - 999.975) <= 5e-4


def test_unesco80_agrees_with_jackett_over_ocean_range():
    worst = 0.0
    for T in np.linspace(-2.0, 35.0, 38):
        for S in np.linspace(0.0, 42.0, 43):
            d = abs(po.rho_unesco80(T, S) - po.rho_from_theta(S, T, 0.0)[0])
            worst = max(worst, d)
    assert worst <= 0.02, worst


def test_unesco80_nonnegsalt():
    assert po.rho_unesco80(10.0, -3.0, nonnegsalt=True) == po.rho_unesco80(10.0, 0.0)
    assert np.isnan(po.rho_unesco80(10.0, -3.0))  # sqrt of a negative S^3


@pytest.mark.parametrize("p", [0.0, 50.0, 1000.0, 5000.0])
def test_eosall_derivatives_are_those_of_rho_from_theta(p):
    rng = np.random.default_rng(int(p))
    h = 1e-4
    for _ in range(40):
        s, th = rng.uniform(1.0, 40.0), rng.uniform(-1.0, 30.0)
        rs, rt = po.eosall_from_theta(s, th, p)
        fs = (po.rho_from_theta(s + h, th, p)[1] - po.rho_from_theta(s - h, th, p)[1]) / (2 * h)
        ft = (po.rho_from_theta(s, th + h, p)[1] - po.rho_from_theta(s, th - h, p)[1]) / (2 * h)
        assert abs(rs - fs) <= 1e-7, (s, th, p, rs, fs)
        assert abs(rt - ft) <= 1e-7, (s, th, p, rt, ft)


def column_domain(kmax=6):
    dom = inputs.make_domain(6, 5, kmax, seed=3, walls=1)
    return dom, po.Oracle.from_domain(dom)


@pytest.mark.parametrize("method", [2, 3])
def test_do_eqstate_columns(method):
    dom, orc = column_domain()
    sh = dom.shape3
    kmax = dom.kmax
    rng = np.random.default_rng(method)
    T = rng.uniform(2.0, 25.0, sh)
    S = rng.uniform(5.0, 36.0, sh)
    hn = rng.uniform(0.5, 4.0, sh)
    land_rho = -777.0
    rho = np.full(sh, land_rho)
    buoy = np.zeros(sh)
    alpha = np.full(sh, 5.0)
    beta = np.full(sh, 5.0)
    assert orc.do_eqstate(method, T, S, hn, rho, buoy, alpha, beta) == 0
    wet = dom.az > 0
    g, rho0 = float(np.float32(9.81)), 1025.0
    for j in range(dom.NY):
        for i in range(dom.NX):
            col = (slice(None), j, i)
            if not wet[j, i]:
                assert np.all(rho[col] == land_rho)
                assert np.all(alpha[col] == 5.0) and np.all(beta[col] == 5.0)
                continue
            for k in range(1, kmax + 1):
                if method == 2:
                    want = po.rho_unesco80(T[k, j, i], S[k, j, i])
                else:
                    want = po.rho_from_theta(S[k, j, i], T[k, j, i], 0.0)[0]
                assert rho[k, j, i] == want
            assert alpha[kmax, j, i] == 0.0 and beta[kmax, j, i] == 0.0
            assert alpha[0, j, i] == 5.0  # below kmin: untouched
            for k in range(1, kmax):
                p = hn[k + 1:, j, i].sum()  # pressure: the layers above interface k
                th = 0.5 * (T[k + 1, j, i] + T[k, j, i])
                s = 0.5 * (S[k + 1, j, i] + S[k, j, i])
                rs, rt = po.eosall_from_theta(s, th, p)
                assert abs(beta[k, j, i] - rs) <= 1e-13 and abs(alpha[k, j, i] - rt) <= 1e-13
    want_buoy = (-g / rho0) * (rho[1:] - rho0)
    assert np.array_equal(buoy[1:], want_buoy)
    assert np.all(buoy[0] == 0.0)


def test_do_eqstate_linear_and_errors():
    dom, orc = column_domain()
    sh = dom.shape3
    rng = np.random.default_rng(1)
    T = rng.uniform(2.0, 25.0, sh)
    S = rng.uniform(5.0, 36.0, sh)
    rho, buoy = np.zeros(sh), np.zeros(sh)
    alpha, beta = np.zeros(sh), np.zeros(sh)
    assert orc.do_eqstate(1, T, S, None, rho, buoy, alpha, beta) == 0
    f32 = lambda v: float(np.float32(v))
    assert np.all(alpha == f32(-0.17)) and np.all(beta == f32(0.78))
    # linear closed form: rho - rho_0 = dtr0 (T - T0) + dsr0 (S - S0)
    want = 1025.0 + f32(-0.17) * (T[1:] - 10.0) + f32(0.78) * (S[1:] - 33.75)
    assert np.allclose(rho[1:], want, rtol=0, atol=1e-12)
    assert orc.do_eqstate(4, T, S, None, rho, buoy) == 1


def test_do_eqstate_nonnegsalt_only_touches_rho():
    dom, orc = column_domain(kmax=3)
    sh = dom.shape3
    T = np.full(sh, 10.0)
    S = np.full(sh, -1.0)
    hn = np.ones(sh)
    rho = np.zeros(sh)
    buoy = np.zeros(sh)
    orc.do_eqstate(3, T, S, hn, rho, buoy, nonnegsalt=True)
    wet = dom.az > 0
    want = po.rho_from_theta(0.0, 10.0, 0.0)[0]
    assert np.all(rho[1:, wet] == want)
