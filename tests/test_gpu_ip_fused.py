"""The equation of state fused into the internal pressure gradients
(gt_eos_internal_pressure, gt_eqstate_internal_pressure, gt_ip_fused.hip)
against gt_eos_linear / gt_do_eqstate followed by gt_ip_blumberg_mellor /
gt_ip_shchepetkin_mcwilliams and against the oracle (eqstate.F90:94-261,
ip_blumberg_mellor.F90:71-152, ip_shchepetkin_mcwilliams.F90:71-288): bit for
bit, every method alone and both together, on open water, on islands with
sentinel-filled outputs (the cells a method does not write keep what they
held) and at full size; UNESCO 1980 and Jackett et al. with and without
NONNEGSALT, the land columns' rho read as do_eqstate leaves it."""
import numpy as np
import pytest
import torch

from code_amd import inputs
from code_amd import transport as gt
from oracle import pyoracle as po
from tests.util import to_dev, to_host

pytestmark = pytest.mark.gpu

CONSTS = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
NAMES = ("zb", "bx", "by", "zs", "sx", "sy")


def with_islands(dom, seed):
    """the domain with land cells scattered through its interior (masks
    derived from az as the reference's do)"""
    rng = np.random.default_rng(seed)
    az = dom.az.copy()
    land = rng.random(az.shape) < 0.12
    az[land] = 0
    au, av, ax = inputs.masks_from_az(az)
    return inputs.Domain(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, dom.H, dom.HU, dom.HV,
                         az, au, av, ax)


def run_case(device, tr, dom, methods, sentinel, sseo):
    """fused GPU, per-method GPU and oracle outputs for one case"""
    T, S = tr.tracers["T"], tr.tracers["S"]
    bm, sm = bool(methods & 1), bool(methods & 2)
    start = {nm: (np.full(dom.shape3, sentinel) if sentinel is not None else np.zeros(dom.shape3))
             for nm in NAMES}
    # the oracle: eos_linear, then the methods asked for
    ref = {nm: a.copy() for nm, a in start.items()}
    orc = po.Oracle.from_domain(dom)
    rho, buoy = np.zeros(dom.shape3), np.zeros(dom.shape3)
    orc.eos_linear(T, S, rho, buoy, *CONSTS)
    if bm:
        orc.ip_blumberg_mellor(dom.H, tr.hn, tr.hun, tr.hvn, buoy, ref["zb"], ref["bx"],
                               ref["by"])
    if sm:
        orc.ip_shchepetkin_mcwilliams(dom.H, tr.hn, tr.hun, tr.hvn, buoy, sseo, ref["zs"],
                                      ref["sx"], ref["sy"])
    ctx = gt.Context.from_domain(dom)
    dv = {nm: to_dev(a, device) for nm, a in (("T", T), ("S", S), ("H", dom.H), ("hn", tr.hn),
                                              ("hun", tr.hun), ("hvn", tr.hvn), ("sseo", sseo))}
    # per-method kernels
    per = {nm: to_dev(a, device) for nm, a in start.items()}
    tb = torch.zeros(dom.shape3, dtype=torch.float64, device=device)
    ctx.eos_linear(dv["T"], dv["S"], None, tb, *CONSTS)
    if bm:
        ctx.ip_blumberg_mellor(dv["H"], dv["hn"], dv["hun"], dv["hvn"], tb, per["zb"], per["bx"],
                               per["by"])
    if sm:
        ctx.ip_shchepetkin_mcwilliams(dv["H"], dv["hn"], dv["hun"], dv["hvn"], tb, dv["sseo"],
                                      per["zs"], per["sx"], per["sy"])
    # fused
    fus = {nm: to_dev(a, device) for nm, a in start.items()}
    ctx.eos_internal_pressure(methods, dv["T"], dv["S"], *CONSTS, dv["H"], dv["hn"], dv["hun"],
                              dv["hvn"], dv["sseo"] if sm else None,
                              *(fus[nm] if on else None for nm, on in zip(
                                  NAMES, (bm, bm, bm, sm, sm, sm))))
    torch.cuda.synchronize()
    ctx.close()
    return ref, {k: to_host(v) for k, v in per.items()}, {k: to_host(v) for k, v in fus.items()}


def check(ref, per, fus, methods):
    used = (["zb", "bx", "by"] if methods & 1 else []) + (["zs", "sx", "sy"] if methods & 2
                                                          else [])
    for nm in NAMES:
        assert np.array_equal(per[nm], ref[nm], equal_nan=True), (nm, "per-method vs oracle")
        assert np.array_equal(fus[nm], ref[nm], equal_nan=True), (nm, "fused vs oracle")
    for nm in used:
        assert np.abs(ref[nm]).max() > 0, nm


def sigma_case(imax, jmax, kmax, seed):
    """T and S on sigma layers (any kmax >= 2)"""
    tr = inputs.config_advection(imax, jmax, kmax, seed, coords="sigma")
    tr.tracers["T"], tr.tracers["S"] = inputs.temperature_salinity(tr, seed)
    return tr


@pytest.mark.parametrize("methods", [1, 2, 3])
@pytest.mark.parametrize("size", [(30, 26, 9), (130, 70, 7), (67, 41, 3), (65, 64, 2)])
def test_fused_open_water(device, size, methods):
    """strips of 60 columns and workgroups of 3 rows, whole and cut short;
    kmax = 2 (the least the methods take) on sigma layers"""
    if size[2] == 2:
        tr = sigma_case(*size, seed=6)
    else:
        tr = inputs.ts_case(seed=5, imax=size[0], jmax=size[1], kmax=size[2])
    sseo = inputs.uniform((tr.dom.NY, tr.dom.NX), 45, -0.2, 0.2)
    check(*run_case(device, tr, tr.dom, methods, None, sseo), methods)


@pytest.mark.parametrize("methods", [1, 2, 3])
def test_fused_islands_sentinels(device, methods):
    """land cells inside the domain, every output pre-filled with the same
    sentinel on both sides: the fused kernels write exactly the cells the
    per-method kernels write"""
    tr = inputs.ts_case(seed=7, imax=90, jmax=61, kmax=11)
    dom = with_islands(tr.dom, 8)
    sseo = inputs.uniform((dom.NY, dom.NX), 46, -0.2, 0.2)
    check(*run_case(device, tr, dom, methods, 3.25, sseo), methods)


@pytest.mark.timeout(300)
def test_fused_full_size(device):
    """400 x 400 x 40 GVC, both methods, against the per-method kernels and
    the oracle"""
    po.set_threads(16)
    tr = inputs.ts_case()
    sseo = inputs.uniform((tr.dom.NY, tr.dom.NX), 47, -0.2, 0.2)
    check(*run_case(device, tr, tr.dom, 3, None, sseo), 3)


def test_fused_rejects(device):
    tr = inputs.ts_case(seed=5, imax=20, jmax=18, kmax=4)
    ctx = gt.Context.from_domain(tr.dom)
    z = torch.zeros(tr.dom.shape3, dtype=torch.float64, device=device)
    H = to_dev(tr.dom.H, device)
    with pytest.raises(gt.GtError):
        ctx.eos_internal_pressure(0, z, z, *CONSTS, H, z, z, z)
    with pytest.raises(ValueError):  # SMcW without sseo
        ctx.eos_internal_pressure(2, z, z, *CONSTS, H, z, z, z, None, None, None, None, z, z, z)
    ctx.close()


@pytest.mark.parametrize("size,kind", [((1, 1, 2), "open"), ((2, 3, 3), "islands"),
                                       ((61, 4, 2), "open"), ((3, 61, 4), "islands"),
                                       ((121, 7, 5), "islands")])
def test_fused_small_and_ragged(device, size, kind):
    """grids of one interior cell, of one strip plus a column, a strip of
    one row group and the like, with land and open edges, random fields"""
    from types import SimpleNamespace
    from tests.test_gpu_parity import random_domain
    dom = random_domain(*size, 50 + size[0], kind)
    sh = dom.shape3
    tr = SimpleNamespace(tracers={"T": inputs.uniform(sh, 61, 5.0, 20.0),
                                  "S": inputs.uniform(sh, 62, 30.0, 35.0)},
                         hn=inputs.uniform(sh, 63, 0.5, 2.0),
                         hun=inputs.uniform(sh, 64, 0.5, 2.0),
                         hvn=inputs.uniform(sh, 65, 0.5, 2.0))
    sseo = inputs.uniform((dom.NY, dom.NX), 66, -0.2, 0.2)
    rho0 = inputs.uniform(sh, 67, 1020.0, 1030.0)
    for methods in (1, 2, 3):
        ref, per, fus = run_case(device, tr, dom, methods, 1.5, sseo)
        for nm in NAMES:
            assert np.array_equal(fus[nm], ref[nm], equal_nan=True), (methods, nm)
            assert np.array_equal(per[nm], ref[nm], equal_nan=True), (methods, nm)
        for eqm in (2, 3):
            ref, per, fus = run_case_eqstate(device, tr, dom, eqm, True, methods, 1.5, sseo,
                                             rho0)
            for nm in NAMES:
                assert np.array_equal(fus[nm], ref[nm], equal_nan=True), (eqm, methods, nm)
                assert np.array_equal(per[nm], ref[nm], equal_nan=True), (eqm, methods, nm)


def run_case_eqstate(device, tr, dom, eqm, nonneg, methods, sentinel, sseo, rho0):
    """do_eqstate of method eqm fused into the pressure gradients: fused GPU,
    do_eqstate + per-method GPU kernels and the oracle's chain, all starting
    from the same rho (kept on land by do_eqstate)"""
    T, S = tr.tracers["T"], tr.tracers["S"]
    bm, sm = bool(methods & 1), bool(methods & 2)
    start = {nm: np.full(dom.shape3, sentinel) for nm in NAMES}
    ref = {nm: a.copy() for nm, a in start.items()}
    orc = po.Oracle.from_domain(dom)
    rho, buoy = rho0.copy(), np.zeros(dom.shape3)
    orc.do_eqstate(eqm, T, S, tr.hn, rho, buoy, nonnegsalt=nonneg)
    if bm:
        orc.ip_blumberg_mellor(dom.H, tr.hn, tr.hun, tr.hvn, buoy, ref["zb"], ref["bx"],
                               ref["by"])
    if sm:
        orc.ip_shchepetkin_mcwilliams(dom.H, tr.hn, tr.hun, tr.hvn, buoy, sseo, ref["zs"],
                                      ref["sx"], ref["sy"])
    ctx = gt.Context.from_domain(dom)
    dv = {nm: to_dev(a, device) for nm, a in (("T", T), ("S", S), ("H", dom.H), ("hn", tr.hn),
                                              ("hun", tr.hun), ("hvn", tr.hvn), ("sseo", sseo),
                                              ("rho", rho0))}
    per = {nm: to_dev(a, device) for nm, a in start.items()}
    trho = to_dev(rho0, device)
    tb = torch.zeros(dom.shape3, dtype=torch.float64, device=device)
    ctx.do_eqstate(eqm, dv["T"], dv["S"], dv["hn"], trho, tb, nonnegsalt=nonneg)
    if bm:
        ctx.ip_blumberg_mellor(dv["H"], dv["hn"], dv["hun"], dv["hvn"], tb, per["zb"], per["bx"],
                               per["by"])
    if sm:
        ctx.ip_shchepetkin_mcwilliams(dv["H"], dv["hn"], dv["hun"], dv["hvn"], tb, dv["sseo"],
                                      per["zs"], per["sx"], per["sy"])
    fus = {nm: to_dev(a, device) for nm, a in start.items()}
    ctx.eqstate_internal_pressure(eqm, methods, dv["T"], dv["S"], dv["rho"], dv["H"], dv["hn"],
                                  dv["hun"], dv["hvn"], dv["sseo"] if sm else None,
                                  *(fus[nm] if on else None for nm, on in zip(
                                      NAMES, (bm, bm, bm, sm, sm, sm))), nonnegsalt=nonneg)
    torch.cuda.synchronize()
    # the fused call leaves rho as it was
    assert torch.equal(dv["rho"], to_dev(rho0, device))
    ctx.close()
    return ref, {k: to_host(v) for k, v in per.items()}, {k: to_host(v) for k, v in fus.items()}


def eqstate_case(seed, nonneg, imax=90, jmax=61, kmax=11):
    """T/S profiles on islands, the land rho unrelated to T and S, and with
    nonneg some salinities below zero"""
    tr = inputs.ts_case(seed=seed, imax=imax, jmax=jmax, kmax=kmax)
    dom = with_islands(tr.dom, seed + 1)
    if nonneg:
        S = tr.tracers["S"].copy()
        neg = np.random.default_rng(seed + 2).random(S.shape) < 0.03
        S[neg] = -0.5
        tr.tracers["S"] = S
    rho0 = inputs.uniform(dom.shape3, seed + 3, 1020.0, 1030.0)
    sseo = inputs.uniform((dom.NY, dom.NX), seed + 4, -0.2, 0.2)
    return tr, dom, rho0, sseo


@pytest.mark.parametrize("methods", [1, 2, 3])
@pytest.mark.parametrize("eqm,nonneg", [(2, False), (2, True), (3, False), (3, True)])
def test_fused_eqstate_islands(device, eqm, nonneg, methods):
    """UNESCO 1980 and Jackett et al. on islands, sentinel-filled outputs:
    bit for bit the chain of do_eqstate and the per-method kernels, and the
    oracle's"""
    tr, dom, rho0, sseo = eqstate_case(20 + eqm, nonneg)
    check(*run_case_eqstate(device, tr, dom, eqm, nonneg, methods, 2.5, sseo, rho0), methods)


def test_fused_eqstate_reads_land_rho(device):
    """Shchepetkin-McWilliams' gradients next to land change with the rho held
    on land and match the chain's for both; Blumberg-Mellor's do not change"""
    tr, dom, rho0, sseo = eqstate_case(31, False)
    rho1 = rho0.copy()
    rho1[:, dom.az == 0] += 1.0
    a = run_case_eqstate(device, tr, dom, 3, False, 3, 0.0, sseo, rho0)
    b = run_case_eqstate(device, tr, dom, 3, False, 3, 0.0, sseo, rho1)
    check(*a, 3)
    check(*b, 3)
    assert not np.array_equal(a[2]["sx"], b[2]["sx"])
    assert not np.array_equal(a[2]["sy"], b[2]["sy"])
    for nm in ("zb", "bx", "by", "zs"):
        assert np.array_equal(a[2][nm], b[2][nm]), nm


@pytest.mark.parametrize("eqm", [2, 3])
def test_fused_eqstate_sigma_kmax2(device, eqm):
    """kmax = 2 (the least the methods take) on sigma layers with islands"""
    tr = sigma_case(65, 64, 2, seed=9)
    dom = with_islands(tr.dom, 10)
    rho0 = inputs.uniform(dom.shape3, 11, 1020.0, 1030.0)
    sseo = inputs.uniform((dom.NY, dom.NX), 12, -0.2, 0.2)
    check(*run_case_eqstate(device, tr, dom, eqm, False, 3, 0.5, sseo, rho0), 3)


@pytest.mark.timeout(300)
def test_fused_eqstate_full_size(device):
    """400 x 400 x 40 GVC, Jackett et al. and both methods"""
    po.set_threads(16)
    tr = inputs.ts_case()
    rho0 = inputs.uniform(tr.dom.shape3, 48, 1020.0, 1030.0)
    sseo = inputs.uniform((tr.dom.NY, tr.dom.NX), 49, -0.2, 0.2)
    check(*run_case_eqstate(device, tr, tr.dom, 3, False, 3, 0.0, sseo, rho0), 3)


def test_fused_eqstate_rejects(device):
    tr = inputs.ts_case(seed=5, imax=20, jmax=18, kmax=4)
    ctx = gt.Context.from_domain(tr.dom)
    z = torch.zeros(tr.dom.shape3, dtype=torch.float64, device=device)
    H = to_dev(tr.dom.H, device)
    for eqm in (0, 4):
        with pytest.raises(gt.GtError, match="not a valid eqstate_method"):
            ctx.eqstate_internal_pressure(eqm, 1, z, z, z, H, z, z, z, None, z, z, z)
    with pytest.raises(ValueError):  # rho is read on land by methods 2 and 3
        ctx.eqstate_internal_pressure(3, 1, z, z, None, H, z, z, z, None, z, z, z)
    # method 1 does without rho
    ctx.eqstate_internal_pressure(1, 1, z, z, None, H, z, z, z, None, z, z, z)
    ctx.close()
