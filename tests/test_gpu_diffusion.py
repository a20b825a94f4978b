"""Vertical diffusion on the GPU against the oracle: salinity bit-exact;
temperature bit-exact without short-wave and within 1e-12 relative with it
(the short-wave profile goes through the device exp, within an ulp of
libm's); the batched tridiagonal solver bit-exact; the 400x400x40 T/S case
at full size with the salt content budget."""
import numpy as np
import pytest
import torch

from code_amd import inputs
from code_amd import transport as gt
from oracle import pyoracle as po
from tests.test_gpu_parity import random_domain
from tests.util import rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12


def fields(dom, seed):
    sh = dom.shape3
    hn = inputs.uniform(sh, seed + 1, 0.2, 3.0)
    nuh = inputs.uniform(sh, seed + 2, 1e-6, 8e-2)
    f = inputs.uniform(sh, seed + 3, -1.0, 30.0)
    return hn, nuh, f


@pytest.mark.parametrize("kind,size", [("islands", (37, 29, 9)), ("open", (130, 70, 2)),
                                       ("walls", (20, 18, 1)), ("walls", (66, 40, 41))])
def test_salinity_diffusion_bit_exact(device, kind, size):
    dom = random_domain(*size, 3, kind)
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    hn, nuh, S = fields(dom, 10)
    dS = to_dev(S, device)
    for cnpar in (0.5, 1.0):
        orc.salinity_diffusion(900.0, cnpar, hn, nuh, S)
        ctx.salinity_diffusion(900.0, cnpar, to_dev(hn, device), to_dev(nuh, device), dS)
        got = to_host(dS)
        assert rel_err(got, S) <= TOL and np.array_equal(got, S), (kind, size, cnpar)


def temp_inputs(dom, seed, swr_max):
    sh2 = dom.az.shape
    return dict(swr=inputs.uniform(sh2, seed, 0.0, swr_max),
                shf=inputs.uniform(sh2, seed + 1, -250.0, 120.0),
                A=inputs.uniform(sh2, seed + 2, 0.5, 0.7),
                g1=inputs.uniform(sh2, seed + 3, 0.3, 1.5),
                g2=inputs.uniform(sh2, seed + 4, 5.0, 20.0),
                dry_z=inputs.uniform(sh2, seed + 5, 0.0, 1.0))


@pytest.mark.parametrize("swr_max,refl", [(0.0, -1.0), (500.0, -1.0), (500.0, 0.4)])
def test_temperature_diffusion(device, swr_max, refl):
    dom = random_domain(53, 31, 7, 5, "islands")
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    hn, nuh, T = fields(dom, 20)
    hn = hn * 0.4  # shallow: the bottom reflection applies where enabled
    S = inputs.uniform(dom.shape3, 7, 20.0, 35.0)
    T[dom.kmax][::3] = -2.5  # some freezing surface cells (sea-ice rule)
    ti = temp_inputs(dom, 40, swr_max)
    rad = np.zeros(dom.shape3)
    drad = to_dev(rad, device)
    dT = to_dev(T, device)
    for cnpar in (0.5, 0.8):
        orc.temperature_diffusion(600.0, cnpar, hn, nuh, *ti.values(), S, rad, T,
                                  swr_bot_refl_frac=refl)
        ctx.temperature_diffusion(600.0, cnpar, to_dev(hn, device), to_dev(nuh, device),
                                  *(to_dev(v, device) for v in ti.values()),
                                  to_dev(S, device), drad, dT, swr_bot_refl_frac=refl)
        got, grad = to_host(dT), to_host(drad)
        assert rel_err(got, T) <= TOL, (swr_max, refl, cnpar, rel_err(got, T))
        assert rel_err(grad, rad) <= TOL
        if swr_max == 0.0:
            assert np.array_equal(got, T) and np.array_equal(grad, rad)


@pytest.mark.parametrize("n,fi,lt", [(1000, 1, 40), (777, 0, 5), (64, 1, 2), (3000, 2, 9)])
def test_tridiagonal_batch_bit_exact(device, n, fi, lt):
    rng = np.random.default_rng(n + lt)
    au = rng.uniform(-1, 0, (lt + 1, n))
    cu = rng.uniform(-1, 0, (lt + 1, n))
    bu = 2.5 + rng.uniform(0, 1, (lt + 1, n))
    du = rng.uniform(-3, 3, (lt + 1, n))
    dom = random_domain(8, 6, 3, 1, "walls")
    ctx = gt.Context.from_domain(dom)
    dv = [to_dev(a, device) for a in (au, bu, cu, du)]
    val = torch.zeros_like(dv[0])
    ctx.tridiagonal_batch(fi, lt, *dv, val)
    got = to_host(val)
    for s in range(0, n, max(1, n // 97)):
        want = po.tridiagonal(lt, fi, lt, au[:, s], bu[:, s], cu[:, s], du[:, s])
        assert np.array_equal(got[fi:, s], want[fi:]), s
    with pytest.raises(gt.GtError):
        ctx.tridiagonal_batch(3, 3, *dv, val)


def test_diffusion_full_size_ts(device):
    """the 400x400x40 T/S case: salinity diffusion bit-exact and its salt
    content unchanged; temperature with short-wave within 1e-12."""
    tr = inputs.ts_case()
    d = tr.dom
    ctx = gt.Context.from_domain(d)
    orc = po.Oracle.from_domain(d)
    sh = d.shape3
    nuh = inputs.uniform(sh, 3, 1e-5, 5e-2)
    T, S = tr.tracers["T"].copy(), tr.tracers["S"].copy()
    dT, dS = to_dev(T, device), to_dev(S, device)
    dhn, dnuh = to_dev(tr.hn, device), to_dev(nuh, device)
    orc.salinity_diffusion(tr.dt, 0.5, tr.hn, nuh, S)
    ctx.salinity_diffusion(tr.dt, 0.5, dhn, dnuh, dS)
    assert np.array_equal(to_host(dS), S)
    wet = d.az == 1
    m0 = (tr.hn[1:] * tr.tracers["S"][1:]).sum(0)[wet]
    m1 = (tr.hn[1:] * S[1:]).sum(0)[wet]
    assert np.allclose(m1, m0, rtol=1e-12, atol=0)
    ti = temp_inputs(d, 90, 600.0)
    ti["dry_z"] = np.ones_like(ti["dry_z"])
    rad = np.zeros(sh)
    drad = to_dev(rad, device)
    orc.temperature_diffusion(tr.dt, 0.5, tr.hn, nuh, *ti.values(), S, rad, T)
    ctx.temperature_diffusion(tr.dt, 0.5, dhn, dnuh, *(to_dev(v, device) for v in ti.values()),
                              dS, drad, dT)
    got = to_host(dT)
    assert rel_err(got, T) <= TOL
    assert rel_err(to_host(drad), rad) <= TOL
