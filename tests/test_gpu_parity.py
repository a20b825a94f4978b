"""FP64 HIP path against the oracle, through the C-ABI.

Tolerance: the bound for the plain FP64 path is 1e-12 relative per
cell; the kernels keep the reference's operation order and are built without
FMA contraction, so they are expected (and reported) to agree bit for bit.
Each test compares full arrays, halos included.
"""
import numpy as np
import pytest

from code_amd import inputs
from code_amd import transport as gt
from oracle import pyoracle as po
from tests.util import rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12
SCHEMES = [gt.UPSTREAM, gt.P2, gt.SUPERBEE, gt.MUSCL, gt.P2_PDM]


def random_domain(imax, jmax, kmax, seed, kind):
    """kind: 'walls' (closed basin), 'open' (water up to and into the halo,
    so the serial halo bounds of the reference are exercised), 'islands'."""
    NX, NY = imax + 4, jmax + 4
    if kind == "walls":
        return inputs.make_domain(imax, jmax, kmax, seed, walls=3)
    az = np.ones((NY, NX), np.int32)
    if kind == "islands":
        u = inputs.uniform((NY, NX), seed + 17)
        az[u < 0.15] = 0
    au, av, ax = inputs.masks_from_az(az)
    H = inputs.uniform((NY, NX), seed + 5, 10.0, 40.0)
    HU = H.copy()
    HV = H.copy()
    HU[:, :-1] = 0.5 * (H[:, :-1] + H[:, 1:])
    HV[:-1, :] = 0.5 * (H[:-1, :] + H[1:, :])
    return inputs.Domain(imax, jmax, kmax, 500.0, 400.0, H, HU, HV, az, au, av, ax)


def random_fields(dom, seed, cmax=0.9, wmax=None, plateaus=True):
    """random tracer, layer heights and transports with |CFL| < cmax"""
    sh = dom.shape3
    f = inputs.uniform(sh, seed + 1)
    if plateaus:
        # a few plateaus so the limiter's deltaf*deltafu<=0 branches are hit
        f[inputs.uniform(sh, seed + 2) < 0.1] = 0.5
    h = inputs.uniform(sh, seed + 3, 0.5, 2.0)
    hu = inputs.uniform(sh, seed + 4, 0.5, 2.0)
    hv = inputs.uniform(sh, seed + 5, 0.5, 2.0)
    dt = 30.0
    uu = inputs.uniform(sh, seed + 6, -1, 1) * hu * cmax * dom.dx / dt
    vv = inputs.uniform(sh, seed + 7, -1, 1) * hv * cmax * dom.dy / dt
    wmax = cmax * 0.5 / dt if wmax is None else wmax
    ww = inputs.uniform(sh, seed + 8, -1, 1) * wmax
    ww[0] = 0.0
    ww[-1] = 0.0
    ho = h.copy()
    hn = inputs.uniform(sh, seed + 9, 0.5, 2.0)
    return dict(f=f, h=h, hu=hu, hv=hv, uu=uu, vv=vv, ww=ww, ho=ho, hn=hn, dt=dt)


@pytest.fixture(scope="module", params=["walls", "open", "islands"])
def case(request, device):
    dom = random_domain(23, 17, 6, 11, request.param)
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    flds = random_fields(dom, 100)
    return dom, ctx, orc, flds, device


def test_reconstruction_bit_exact(device):
    import torch
    dom = random_domain(8, 8, 2, 1, "walls")
    ctx = gt.Context.from_domain(dom)
    n = 20000
    cfl = inputs.uniform((n,), 1, 0.0, 1.0)
    cfl[:50] = [0.0, 1.0, 0.5, 1e-12, 0.999999] * 10
    fuu = inputs.uniform((n,), 2, -1, 1)
    fu = inputs.uniform((n,), 3, -1, 1)
    fd = inputs.uniform((n,), 4, -1, 1)
    fd[:100] = fu[:100]        # deltaf = 0
    fuu[100:200] = fu[100:200]  # deltafu = 0
    for scheme in SCHEMES:
        out = torch.empty(n, dtype=torch.float64, device=device)
        ctx.adv_interfacial_reconstruction(scheme, *(to_dev(a, device) for a in (cfl, fuu, fu, fd)),
                                           out)
        ref = np.array([po.reconstruct(scheme, *v) for v in zip(cfl, fuu, fu, fd)])
        got = to_host(out)
        assert np.array_equal(np.isfinite(got), np.isfinite(ref))
        fin = np.isfinite(ref)
        assert rel_err(got[fin], ref[fin]) <= TOL, scheme
        assert np.array_equal(got[fin], ref[fin]), f"scheme {scheme} not bit-exact"


@pytest.mark.parametrize("direction", ["u", "v"])
def test_split_uv(case, direction):
    dom, ctx, orc, F, dev = case
    k = 3
    for scheme in SCHEMES:
        for AH in (-1.0, 7.5):
            for tag in (gt.H_TAG, gt.U_TAG, gt.V_TAG):
                for alias in (True, False):
                    f = F["f"][k].copy()
                    fi = f if alias else F["f"][k - 1].copy()
                    Di = F["h"][k].copy()
                    adv = F["hn"][k].copy()
                    vel = F["uu"][k] if direction == "u" else F["vv"][k]
                    DV = F["hu"][k] if direction == "u" else F["hv"][k]
                    tf, tfi, tDi, tadv = (to_dev(a, dev) for a in (f, fi, Di, adv))
                    if alias:
                        tfi = tf
                    tvel, tDV = to_dev(vel, dev), to_dev(DV, dev)
                    fn = ctx.adv_split_u if direction == "u" else ctx.adv_split_v
                    fn(F["dt"], tf, tfi, tDi, tadv, tvel, tDV, 0.5, scheme, AH, tag)
                    on = orc.adv_split_u if direction == "u" else orc.adv_split_v
                    on(F["dt"], f, fi, Di, adv, vel, DV, 0.5, scheme, AH, tag)
                    for got, ref, nm in ((tfi, fi, "fi"), (tDi, Di, "Di"), (tadv, adv, "adv")):
                        e = rel_err(to_host(got), ref)
                        assert e <= TOL, (direction, scheme, AH, tag, alias, nm, e)


@pytest.mark.parametrize("itersmax", [1, 4])
def test_split_w(case, itersmax):
    dom, ctx, orc, F, dev = case
    dt = F["dt"]
    # vertical CFL up to ~2.5 so that sub-cycling and its restarts happen
    ww = F["ww"] * (5.0 if itersmax > 1 else 1.0)
    for scheme in SCHEMES:
        for tag in (gt.H_TAG, gt.W_TAG):
            for alias in (True, False):
                f = F["f"].copy()
                fi = f if alias else F["hn"].copy()
                hi = F["h"].copy()
                adv = F["uu"].copy()
                tf = to_dev(f, dev)
                tfi = tf if alias else to_dev(fi, dev)
                thi, tadv = to_dev(hi, dev), to_dev(adv, dev)
                ctx.warnings()
                ctx.adv_split_w(dt, tf, tfi, thi, tadv, to_dev(ww, dev), 1.0, scheme, tag,
                                itersmax)
                nw_gpu = ctx.warnings()
                nw = orc.adv_split_w(dt, f, fi, hi, adv, ww, 1.0, scheme, tag, itersmax)
                assert nw_gpu == nw
                for got, ref, nm in ((tfi, fi, "fi"), (thi, hi, "hi"), (tadv, adv, "adv")):
                    e = rel_err(to_host(got), ref)
                    assert e <= TOL, (itersmax, scheme, tag, alias, nm, e)


def test_upstream_2dh(case):
    dom, ctx, orc, F, dev = case
    k = 2
    for AH in (-1.0, 4.0):
        for alias in (True, False):
            f = F["f"][k].copy()
            fi = f if alias else F["f"][k + 1].copy()
            Di, adv = F["h"][k].copy(), F["ww"][k].copy()
            args = [F["uu"][k], F["vv"][k], F["hn"][k], F["hu"][k], F["hv"][k]]
            tf = to_dev(f, dev)
            tfi = tf if alias else to_dev(fi, dev)
            tDi, tadv = to_dev(Di, dev), to_dev(adv, dev)
            ctx.adv_upstream_2dh(F["dt"], tf, tfi, tDi, tadv, *(to_dev(a, dev) for a in args), AH)
            orc.adv_upstream_2dh(F["dt"], f, fi, Di, adv, *args, AH)
            for got, ref in ((tfi, fi), (tDi, Di), (tadv, adv)):
                assert rel_err(to_host(got), ref) <= TOL


def test_do_advection_2d(case):
    dom, ctx, orc, F, dev = case
    k = 4
    for split in (gt.NOSPLIT, gt.FULLSPLIT, gt.HALFSPLIT, gt.HVSPLIT):
        for scheme in [gt.NOADV, gt.UPSTREAM_2DH] + SCHEMES + [gt.J7]:
            for outs in (True, False):
                f = F["f"][k].copy()
                args = [F["uu"][k], F["vv"][k], F["hu"][k], F["hv"][k], F["ho"][k], F["hn"][k]]
                Dires = np.zeros_like(f) if outs else None
                advres = np.zeros_like(f) if outs else None
                rc = orc.do_advection(F["dt"], f, *args, split, scheme, 2.0, gt.H_TAG, Dires,
                                      advres)
                tf = to_dev(F["f"][k], dev)
                tD = to_dev(np.zeros_like(f), dev) if outs else None
                tA = to_dev(np.zeros_like(f), dev) if outs else None
                try:
                    ctx.do_advection(F["dt"], tf, *(to_dev(a, dev) for a in args), split, scheme,
                                     2.0, gt.H_TAG, tD, tA)
                    code = 0
                except gt.GtError as e:
                    code = e.code
                if rc != 0 or code != 0:
                    assert code != 0 and rc != 0, (split, scheme, rc, code)
                    continue
                assert rel_err(to_host(tf), f) <= TOL, (split, scheme)
                if outs:
                    assert rel_err(to_host(tD), Dires) <= TOL
                    assert rel_err(to_host(tA), advres) <= TOL


HS = [gt.NOADV, gt.UPSTREAM, gt.UPSTREAM_2DH, gt.P2, gt.SUPERBEE, gt.MUSCL, gt.P2_PDM]
VS = [gt.NOADV, gt.UPSTREAM, gt.P2, gt.SUPERBEE, gt.MUSCL, gt.P2_PDM]


@pytest.mark.parametrize("split", [gt.NOSPLIT, gt.FULLSPLIT, gt.HALFSPLIT, gt.HVSPLIT])
def test_do_advection_3d_all_schemes(case, split):
    dom, ctx, orc, F, dev = case
    names = ["uu", "vv", "ww", "hu", "hv", "ho", "hn"]
    dargs = [to_dev(F[n], dev) for n in names]
    for itv in (1, 3):
        ctx.set_adv_ver_iterations(itv)
        for hs in HS:
            for vs in VS:
                for tag in (gt.H_TAG, gt.W_TAG, gt.U_TAG):
                    f = F["f"].copy()
                    hires = np.zeros_like(f)
                    advres = np.zeros_like(f)
                    rc, _ = orc.do_advection_3d(F["dt"], f, *(F[n] for n in names), split, hs, vs,
                                                1.5, tag, itv, hires, advres)
                    tf = to_dev(F["f"], dev)
                    th = to_dev(np.zeros_like(f), dev)
                    ta = to_dev(np.zeros_like(f), dev)
                    try:
                        ctx.do_advection_3d(F["dt"], tf, *dargs, split, hs, vs, 1.5, tag, th, ta)
                        code = 0
                    except gt.GtError as e:
                        code = e.code
                    if rc != 0 or code != 0:
                        assert rc != 0 and code != 0, (split, hs, vs, tag, rc, code)
                        continue
                    for got, ref, nm in ((tf, f, "f"), (th, hires, "hires"), (ta, advres, "advres")):
                        e = rel_err(to_host(got), ref)
                        assert e <= TOL, (split, hs, vs, tag, itv, nm, e)
    ctx.set_adv_ver_iterations(1)


def test_config1_trajectory_bit_exact(device):
    """The Gaussian-blob case at full size: 100x100x20 sigma, HALFSPLIT
    Superbee, 100 steps."""
    tr = inputs.blob_case()
    d = tr.dom
    ctx = gt.Context.from_domain(d)
    orc = po.Oracle.from_domain(d)
    f = tr.tracers["c"].copy()
    m0, v0 = orc.budget(tr.hn, f)
    names = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn]
    dn = [to_dev(a, device) for a in names]
    tf = to_dev(f, device)
    for _ in range(100):
        orc.do_advection_3d(tr.dt, f, *names, gt.HALFSPLIT, gt.SUPERBEE, gt.SUPERBEE, -1.0,
                            gt.H_TAG)
        ctx.do_advection_3d(tr.dt, tf, *dn, gt.HALFSPLIT, gt.SUPERBEE, gt.SUPERBEE, -1.0, gt.H_TAG)
    got = to_host(tf)
    assert rel_err(got, f) <= TOL
    assert np.array_equal(got, f), "config-1 trajectory not bit-exact"
    m1, _ = ctx.budget(to_dev(tr.hn, device), tf)
    assert abs(m1 - m0) <= 1e-13 * abs(m0)


def ip_fields(seed=2, imax=30, jmax=26, kmax=9):
    tr = inputs.ts_case(seed=seed, imax=imax, jmax=jmax, kmax=kmax)
    d = tr.dom
    T, S = tr.tracers["T"], tr.tracers["S"]
    sseo = inputs.uniform((d.NY, d.NX), seed + 40, -0.2, 0.2)
    return tr, d, T, S, sseo


@pytest.mark.parametrize("size", [(30, 26, 9), (130, 70, 7)])
def test_eos_and_internal_pressure(device, size):
    """one strip and several strips (60 columns x 16 rows) of the SMcW
    kernel; bit-exact against the oracle"""
    import torch
    tr, d, T, S, sseo = ip_fields(imax=size[0], jmax=size[1], kmax=size[2])
    ctx = gt.Context.from_domain(d)
    orc = po.Oracle.from_domain(d)
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    rho = np.zeros(d.shape3)
    buoy = np.zeros(d.shape3)
    orc.eos_linear(T, S, rho, buoy, *consts)
    trho = torch.zeros(d.shape3, dtype=torch.float64, device=device)
    tbuoy = torch.zeros_like(trho)
    ctx.eos_linear(to_dev(T, device), to_dev(S, device), trho, tbuoy, *consts)
    assert rel_err(to_host(trho), rho) <= TOL
    assert rel_err(to_host(tbuoy), buoy) <= TOL
    H = d.H
    for method in ("bm", "sm"):
        zz = np.zeros(d.shape3)
        idx = np.zeros(d.shape3)
        idy = np.zeros(d.shape3)
        tz, tx, ty = (torch.zeros_like(trho) for _ in range(3))
        dargs = [to_dev(a, device) for a in (H, tr.hn, tr.hun, tr.hvn, buoy)]
        if method == "bm":
            orc.ip_blumberg_mellor(H, tr.hn, tr.hun, tr.hvn, buoy, zz, idx, idy)
            ctx.ip_blumberg_mellor(*dargs, tz, tx, ty)
        else:
            orc.ip_shchepetkin_mcwilliams(H, tr.hn, tr.hun, tr.hvn, buoy, sseo, zz, idx, idy)
            ctx.ip_shchepetkin_mcwilliams(*dargs, to_dev(sseo, device), tz, tx, ty)
        assert np.abs(idx).max() > 0 and np.abs(idy).max() > 0
        for got, ref, nm in ((tz, zz, "zz"), (tx, idx, "idpdx"), (ty, idy, "idpdy")):
            e = rel_err(to_host(got), ref)
            assert e <= TOL, (method, nm, e)
            assert np.array_equal(to_host(got), ref), (method, nm, "not bit-exact")


def test_halo_fill(case):
    import torch
    dom, ctx, orc, F, dev = case
    for alias in (True, False):
        f1 = F["f"].copy()
        f2 = f1 if alias else F["h"].copy()
        t1 = to_dev(f1, dev)
        t2 = t1 if alias else to_dev(f2, dev)
        orc.update_3d_halo(f1, f2, 1)
        ctx.update_3d_halo(t1, t2, gt.H_TAG, 1)
        assert np.array_equal(to_host(t1), f1)
    # 2-D: the SAVE'd mirror flag carries over to calls without `mirror`
    for alias in (True, False):
        state = 1
        for mirror in (1, -1, 0, -1, 1, -1):
            f1 = F["f"][2].copy()
            f2 = f1 if alias else F["h"][2].copy()
            t1 = to_dev(f1, dev)
            t2 = t1 if alias else to_dev(f2, dev)
            state = orc.update_2d_halo(f1, f2, mirror, state)
            ctx.update_2d_halo(t1, t2, gt.H_TAG, mirror)
            assert np.array_equal(to_host(t1), f1), (mirror, alias)


def test_tile_exchange(device):
    px, py, nk = 3, 2, 4
    ni, nj = [], []
    for ty in range(py):
        for tx in range(px):
            ni.append([7, 5, 9][tx])
            nj.append([6, 8][ty])
    tiles = [inputs.uniform((nk, nj[t] + 4, ni[t] + 4), 300 + t) for t in range(px * py)]
    ref = [t.copy() for t in tiles]
    po.exchange_tiles(px, py, ni, nj, nk, ref)
    dt = [to_dev(t, device) for t in tiles]
    gt.exchange_tiles(px, py, ni, nj, nk, dt)
    for t in range(px * py):
        assert np.array_equal(to_host(dt[t]), ref[t]), t


def test_budget(case):
    dom, ctx, orc, F, dev = case
    m, v = orc.budget(F["h"], F["f"])
    gm, gv = ctx.budget(to_dev(F["h"], dev), to_dev(F["f"], dev))
    assert abs(gm - m) <= 1e-15 * abs(m)
    assert abs(gv - v) <= 1e-15 * abs(v)


@pytest.mark.parametrize("nt", [1, 6, 7, 20])
def test_budget_batch_f32_equals_per_tracer(device, nt):
    """gt_budget_batch_f32 (one pass, groups of six tracers, the last one
    partial for 7 and 20) equals gt_budget_f32 per tracer bit for bit, and
    the oracle's double-double sum of the FP32 values to 1e-15"""
    import torch
    dom = random_domain(37, 29, 6, 11, "islands")
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    rng = np.random.default_rng(nt)
    h = rng.uniform(0.5, 3.0, dom.shape3)
    f = rng.uniform(-2.0, 5.0, (nt,) + dom.shape3).astype(np.float32)
    th, tf = to_dev(h, device), torch.from_numpy(f).to(device)
    out = ctx.budget_batch(th, tf).cpu().numpy()
    for q in range(nt):
        m, v = ctx.budget(th, tf[q])
        assert out[q] == m, (q, out[q], m)
        assert out[nt] == v
        mo, vo = orc.budget(h, f[q].astype(np.float64))
        assert abs(out[q] - mo) <= 1e-15 * abs(mo)
        assert abs(out[nt] - vo) <= 1e-15 * abs(vo)
    with pytest.raises(ValueError):
        ctx.budget_batch(th, tf, ntracers=nt + 1)


def test_config2_full_size_properties(device):
    """The T/S case at full size (400x400x40 GVC, P2-PDM): one step conserves tracer
    mass to round-off, keeps a constant field constant, and ends the volume
    trajectory on hn."""
    import torch
    tr = inputs.ts_case()
    d = tr.dom
    ctx = gt.Context.from_domain(d)
    dn = [to_dev(a, device) for a in (tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn)]
    tS = to_dev(tr.tracers["S"], device)
    th = torch.zeros_like(tS)
    ta = torch.zeros_like(tS)
    thn = dn[-1]
    m0, v0 = ctx.budget(thn, tS)
    ctx.do_advection_3d(tr.dt, tS, *dn, gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM, -1.0, gt.H_TAG, th, ta)
    m1, v1 = ctx.budget(th, tS)
    assert abs(m1 - m0) <= 1e-12 * abs(m0), (m0, m1)
    wet = d.az == 1
    hires = to_host(th)
    assert np.abs(hires[1:, wet] - tr.hn[1:, wet]).max() <= 1e-9
    c = torch.full_like(tS, 2.5)
    ctx.do_advection_3d(tr.dt, c, *dn, gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM, -1.0, gt.H_TAG)
    cc = to_host(c)
    assert np.abs(cc[1:, wet] - 2.5).max() <= 1e-12


@pytest.mark.parametrize("substr", [False, True])
def test_do_internal_pressure_dispatch(device, substr):
    """gt_do_internal_pressure (internal_pressure.F90:181-274): methods 1 and 6
    through the dispatch with the context's own zz, with and without
    SUBSTR_INI_PRESS (:235-258), over three calls whose buoyancy changes;
    bit-exact against the oracle (outputs start from the same values, since
    the routines leave cells outside their loops alone).  Methods outside
    this build give GT_ENOTIMPL, unknown ones 'Not valid ip_method
    specified'."""
    import torch
    tr, d, T, S, sseo = ip_fields(imax=70, jmax=40, kmax=8)
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    dH, dhn, dhun, dhvn = (to_dev(a, device) for a in (d.H, tr.hn, tr.hun, tr.hvn))
    dsse = to_dev(sseo, device)
    for method in (gt.BLUMBERG_MELLOR, gt.SHCHEPETKIN_MCWILLIAMS):
        ctx = gt.Context.from_domain(d)
        orc = po.Oracle.from_domain(d)
        if substr:
            ctx.set_substr_ini_press(True)
        state = {} if substr else None
        zz = np.zeros(d.shape3)
        for q in range(3):
            Tq = np.where(T > -9000, T + 0.7 * q * inputs.uniform(T.shape, 500 + q, -1, 1), T)
            rho, buoy = np.zeros(d.shape3), np.zeros(d.shape3)
            orc.eos_linear(Tq, S, rho, buoy, *consts)
            ix, iy = np.full(d.shape3, 7.0), np.full(d.shape3, 7.0)
            orc.do_internal_pressure(method, d.H, tr.hn, tr.hun, tr.hvn, buoy, sseo, zz, ix, iy,
                                     substr=state)
            tx = torch.full(d.shape3, 7.0, dtype=torch.float64, device=device)
            ty = torch.full_like(tx, 7.0)
            ctx.do_internal_pressure(method, dH, dhn, dhun, dhvn, to_dev(buoy, device),
                                     dsse, tx, ty)
            assert np.array_equal(to_host(tx), ix), (method, substr, q)
            assert np.array_equal(to_host(ty), iy), (method, substr, q)
            if substr and q == 0:
                assert np.all(ix[:, 2:-2, 2:-2] == 0.0)
            if q > 0:
                assert np.abs(ix[1:, 2:-2, 2:-2]).max() > 0
    ctx = gt.Context.from_domain(d)
    z = torch.zeros(d.shape3, dtype=torch.float64, device=device)
    for m, code in ((2, 8), (3, 8), (4, 8), (5, 8), (7, 8), (0, 3), (8, 3), (-1, 3)):
        with pytest.raises(gt.GtError) as e:
            ctx.do_internal_pressure(m, dH, dhn, dhun, dhvn, z, dsse, z, z)
        assert e.value.code == code, (m, e.value.code)


def test_context_derives_uvx_masks_from_az(device):
    """gt_create with au/av/ax NULL derives them as domain.F90:294-346 does;
    the step then equals the oracle's (masks given) bit for bit, on a basin
    with islands and open-boundary cells (az = 2)"""
    dom = random_domain(41, 23, 6, 13, "islands")
    az = dom.az.copy()
    az[5:9, 2] = 2
    au, av, ax = inputs.masks_from_az(az)
    d = inputs.Domain(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, dom.H, dom.HU, dom.HV, az,
                      au, av, ax)
    F = random_fields(d, 700)
    ctx = gt.Context(d.imax, d.jmax, d.kmax, d.dx, d.dy, az)       # masks derived in C
    orc = po.Oracle.from_domain(d)
    host = [F[k] for k in ("uu", "vv", "ww", "hu", "hv", "ho", "hn")]
    dv = [to_dev(a, device) for a in host]
    for split, hs, vs in ((gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM), (gt.NOSPLIT, gt.SUPERBEE,
                                                                   gt.MUSCL)):
        for tag in (gt.H_TAG, gt.U_TAG, gt.V_TAG):
            f = F["f"].copy()
            tf = to_dev(f, device)
            orc.do_advection_3d(F["dt"], f, *host, split, hs, vs, -1.0, tag)
            ctx.do_advection_3d(F["dt"], tf, *dv, split, hs, vs, -1.0, tag)
            assert np.array_equal(to_host(tf), f), (split, tag)


def test_batch_rejects_a_shared_context(device):
    """two batch entries on one context would share its work fields: the
    C-ABI refuses with GT_EARG (the Python binding refuses earlier)"""
    import ctypes as C
    import torch
    dom = random_domain(20, 12, 4, 3, "walls")
    c1, c2 = gt.Context.from_domain(dom), gt.Context.from_domain(dom)
    F = random_fields(dom, 40)
    dv = {k: to_dev(F[k], device) for k in ("uu", "vv", "ww", "hu", "hv", "ho")}
    mk = lambda c: dict(ctx=c, f_in=torch.zeros((2,) + dom.shape3, dtype=torch.float32,
                                                device=device),
                        f_out=torch.zeros((2,) + dom.shape3, dtype=torch.float32,
                                          device=device), **dv)
    doms = [mk(c1), mk(c2)]
    arr = gt._batch_array(doms, 2)
    arr[1].ctx = c1.h.value
    lib = gt.load_library()
    rc = lib.gt_adv3d_batch_f32(2, arr, 2, 30.0, gt.P2_PDM, gt.P2_PDM, -1.0, gt.H_TAG, None)
    assert rc == 7, rc
    assert b"share one context" in lib.gt_last_error()
    with pytest.raises(ValueError):
        gt.adv3d_batch([mk(c1), mk(c1)], 2, 30.0, gt.P2_PDM, gt.P2_PDM)


def test_reserve_sizes_the_work_fields(device):
    """after gt_reserve the step runs as before (bit-exact with the oracle)"""
    dom = random_domain(33, 21, 6, 5, "open")
    F = random_fields(dom, 900)
    ctx = gt.Context.from_domain(dom)
    ctx.set_adv_ver_iterations(3)
    ctx.reserve(18, 2)
    orc = po.Oracle.from_domain(dom)
    host = [F[k] for k in ("uu", "vv", "ww", "hu", "hv", "ho", "hn")]
    dv = [to_dev(a, device) for a in host]
    for split in (gt.HALFSPLIT, gt.NOSPLIT, gt.FULLSPLIT, gt.HVSPLIT):
        f = F["f"].copy()
        tf = to_dev(f, device)
        orc.do_advection_3d(F["dt"], f, *host, split, gt.SUPERBEE, gt.SUPERBEE, -1.0, gt.H_TAG,
                            adv_ver_iterations=3)
        ctx.do_advection_3d(F["dt"], tf, *dv, split, gt.SUPERBEE, gt.SUPERBEE, -1.0, gt.H_TAG)
        assert np.array_equal(to_host(tf), f), split


@pytest.mark.timeout(900)
def test_config2_full_size_200_steps_against_oracle(device):
    """The T/S workload at full size (400x400x40 GVC) for the config's 200
    steps, on the path the benchmark times: the fused EOS + Blumberg-Mellor +
    Shchepetkin-McWilliams launch pair and the batched FP64 P2-PDM HALFSPLIT
    advection of T and S in one launch pair (default strip tuning), against
    the oracle's eos_linear, ip_* and do_advection_3d per tracer: T, S and
    the four pressure gradients of the last step bit for bit"""
    import torch
    po.set_threads(16)
    tr = inputs.ts_case()
    d = tr.dom
    sseo = inputs.uniform((d.NY, d.NX), 48, -0.1, 0.1)
    ctx = gt.Context.from_domain(d)
    orc = po.Oracle.from_domain(d)
    host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn]
    uu, vv, ww, hun, hvn, ho, hn = (to_dev(a, device) for a in host)
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    T, S = tr.tracers["T"].copy(), tr.tracers["S"].copy()
    TS = to_dev(np.stack([T, S]), device)
    TS2 = torch.empty_like(TS)
    z3 = lambda: np.zeros(d.shape3)  # noqa: E731
    rho, buoy = z3(), z3()
    out = {k: z3() for k in ("zb", "bx", "by", "zs", "sx", "sy")}
    tout = {k: torch.zeros(d.shape3, dtype=torch.float64, device=device) for k in out}
    dH, dsse = to_dev(d.H, device), to_dev(sseo, device)
    for step in range(200):
        orc.eos_linear(T, S, rho, buoy, *consts)
        orc.ip_blumberg_mellor(d.H, tr.hn, tr.hun, tr.hvn, buoy, out["zb"], out["bx"], out["by"])
        orc.ip_shchepetkin_mcwilliams(d.H, tr.hn, tr.hun, tr.hvn, buoy, sseo, out["zs"],
                                      out["sx"], out["sy"])
        for f in (T, S):
            orc.do_advection_3d(tr.dt, f, *host, gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM, -1.0,
                                gt.H_TAG)
        ctx.eos_internal_pressure(3, TS[0], TS[1], *consts, dH, hn, hun, hvn, dsse,
                                  *(tout[k] for k in ("zb", "bx", "by", "zs", "sx", "sy")))
        gt.adv3d_batch([dict(ctx=ctx, f_in=TS, f_out=TS2, uu=uu, vv=vv, ww=ww, hu=hun, hv=hvn,
                             ho=ho)], 2, tr.dt, gt.P2_PDM, gt.P2_PDM)
        TS, TS2 = TS2, TS
        if step % 50 == 49:
            print(f"step {step + 1}", flush=True)
    torch.cuda.synchronize()
    got = {"T": TS[0], "S": TS[1], **tout}
    want = {"T": T, "S": S, **out}
    for nm in want:
        h = to_host(got[nm])
        assert rel_err(h, want[nm]) <= TOL, (nm, rel_err(h, want[nm]))
        assert np.array_equal(h, want[nm]), (nm, "not bit-exact")
