"""The full equation of state on the GPU (gt_do_eqstate and the point
functions) against the oracle: bit-exact (bound written: 1e-12 relative,
asserted as array_equal), every method, NONNEGSALT on and off, with and
without the expansion coefficients, land columns keeping their rho; the
reference's own check values evaluated on the device; the 400x400x40 T/S
case at full size."""
import numpy as np
import pytest
import torch

from code_amd import inputs
from code_amd import transport as gt
from oracle import pyoracle as po
from tests.test_gpu_parity import random_domain
from tests.test_oracle_eqstate import GOLDEN
from tests.util import rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_point_functions_bit_exact(device):
    dom = random_domain(8, 6, 3, 1, "walls")
    ctx = gt.Context.from_domain(dom)
    rng = np.random.default_rng(11)
    n = 4000
    s = rng.uniform(0.0, 42.0, n)
    th = rng.uniform(-2.0, 32.0, n)
    p = rng.uniform(0.0, 6000.0, n)
    p[::7] = 0.0
    g1 = GOLDEN["rho_from_theta"][0]
    s[0], th[0], p[0] = g1["s"], g1["th"], g1["p"]  # the reference's check value
    ds, dth, dp = (to_dev(a, device) for a in (s, th, p))
    o1 = torch.empty_like(ds)
    o2 = torch.empty_like(ds)
    ctx.rho_from_theta(ds, dth, dp, o1, o2)
    got0, gotp = to_host(o1), to_host(o2)
    want = np.array([po.rho_from_theta(*a) for a in zip(s, th, p)])
    assert np.array_equal(got0, want[:, 0]) and np.array_equal(gotp, want[:, 1])
    assert abs(gotp[0] - g1["densp"]) <= g1["tol_rel"] * g1["densp"]
    ctx.eosall_from_theta(ds, dth, dp, o1, o2)
    want = np.array([po.eosall_from_theta(*a) for a in zip(s, th, p)])
    assert np.array_equal(to_host(o1), want[:, 0]) and np.array_equal(to_host(o2), want[:, 1])
    g2 = GOLDEN["eosall_from_theta"][0]
    assert abs(to_host(o1)[0] - g2["rho_s"]) <= g2["tol_abs"]
    assert abs(to_host(o2)[0] - g2["rho_th"]) <= g2["tol_abs"]
    for nonneg in (False, True):
        s2 = s.copy()
        s2[::5] = -rng.uniform(0.0, 2.0, len(s2[::5]))
        ds2 = to_dev(s2, device)
        ctx.rho_from_theta_unesco80(dth, ds2, o1, nonnegsalt=nonneg)
        want = np.array([po.rho_unesco80(t, v, nonneg) for t, v in zip(th, s2)])
        got = to_host(o1)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok]), rel_err(got[ok], want[ok])
    # p omitted = 0 dbar
    ctx.rho_from_theta(ds, dth, None, o1, None)
    want = np.array([po.rho_from_theta(a, b, 0.0)[0] for a, b in zip(s, th)])
    assert np.array_equal(to_host(o1), want)


@pytest.mark.parametrize("kind,size", [("islands", (37, 29, 7)), ("open", (130, 70, 9)),
                                       ("walls", (20, 18, 1))])
def test_do_eqstate_bit_exact(device, kind, size):
    dom = random_domain(*size, 5, kind)
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    sh = dom.shape3
    T = inputs.uniform(sh, 21, -1.0, 28.0)
    S = inputs.uniform(sh, 22, -0.5, 38.0)  # a few negative salinities
    hn = inputs.uniform(sh, 23, 0.3, 5.0)
    dT, dS, dh = (to_dev(a, device) for a in (T, S, hn))
    for method in (1, 2, 3):
        for nonneg in (False, True):
            for ab in (True, False):
                rho0 = inputs.uniform(sh, 24, 1000.0, 1030.0)  # land keeps these
                rho, buoy = rho0.copy(), np.zeros(sh)
                alpha = np.full(sh, 3.0) if ab else None
                beta = np.full(sh, 4.0) if ab else None
                assert orc.do_eqstate(method, T, S, hn, rho, buoy, alpha, beta,
                                      nonnegsalt=nonneg) == 0
                drho, dbuoy = to_dev(rho0, device), torch.zeros(sh, dtype=torch.float64,
                                                                device=device)
                da = to_dev(np.full(sh, 3.0), device) if ab else None
                db = to_dev(np.full(sh, 4.0), device) if ab else None
                ctx.do_eqstate(method, dT, dS, dh, drho, dbuoy, da, db, nonnegsalt=nonneg)
                case = (kind, method, nonneg, ab)
                for name, g, w in (("rho", drho, rho), ("buoy", dbuoy, buoy), ("alpha", da, alpha),
                                   ("beta", db, beta)):
                    if w is None:
                        continue
                    gh = to_host(g)
                    same = np.array_equal(gh, w, equal_nan=True)
                    assert same, (case, name, rel_err(gh, w))


def test_do_eqstate_errors(device):
    dom = random_domain(8, 6, 3, 1, "walls")
    ctx = gt.Context.from_domain(dom)
    z = torch.zeros(dom.shape3, dtype=torch.float64, device=device)
    with pytest.raises(gt.GtError, match="eqstate_method"):
        ctx.do_eqstate(4, z, z, z, z, z)
    with pytest.raises(gt.GtError):
        ctx.do_eqstate(3, z, z, None, z, z, z, z)  # alpha/beta need hn


def test_do_eqstate_full_size_ts(device):
    """Jackett EOS with expansion coefficients on the 400x400x40 GVC T/S
    case, against the oracle, bit for bit."""
    tr = inputs.ts_case()
    d = tr.dom
    T, S = tr.tracers["T"], tr.tracers["S"]
    ctx = gt.Context.from_domain(d)
    orc = po.Oracle.from_domain(d)
    sh = d.shape3
    rho, buoy, alpha, beta = (np.zeros(sh) for _ in range(4))
    orc.do_eqstate(3, T, S, tr.hn, rho, buoy, alpha, beta)
    dv = [to_dev(a, device) for a in (T, S, tr.hn)]
    out = [torch.zeros(sh, dtype=torch.float64, device=device) for _ in range(4)]
    ctx.do_eqstate(3, *dv, *out)
    for name, g, w in zip(("rho", "buoy", "alpha", "beta"), out, (rho, buoy, alpha, beta)):
        gh = to_host(g)
        assert rel_err(gh, w) <= TOL, name
        assert np.array_equal(gh, w), name
    wet = d.az == 1
    # stable stratification of the case: alpha < 0 (warmer above), beta > 0
    assert np.all(alpha[1:d.kmax, wet] < 0) and np.all(beta[1:d.kmax, wet] > 0)
