"""The HALFSPLIT strip kernels and the batched multi-tracer step.

* FP64 (strip kernels) must equal the per-sub-step kernels and the oracle bit
  for bit (bound written: 1e-12 relative per cell), for every scheme pair,
  tag, strip height, with and without the optional outputs, on domains
  narrower and wider than one strip.
* FP32 batched: max and RMS error against the FP64 oracle, normalised by the
  field's range, within 1e-5 after several steps; the FP64 mass budget of
  each tracer is conserved to the FP32 rounding of the stored tracer.
"""
import numpy as np
import pytest
import torch

from code_amd import inputs
from code_amd import transport as gt
from oracle import pyoracle as po
from tests.test_gpu_parity import random_domain, random_fields
from tests.util import rel_err, to_dev, to_host

pytestmark = pytest.mark.gpu
TOL = 1e-12
SCHEMES = [gt.UPSTREAM, gt.P2, gt.SUPERBEE, gt.MUSCL, gt.P2_PDM]
NAMES = ["uu", "vv", "ww", "hu", "hv", "ho", "hn"]


@pytest.mark.parametrize("kind", ["walls", "open", "islands"])
def test_fused_f64_bit_exact(device, kind):
    dom = random_domain(37, 29, 7, 5, kind)
    F = random_fields(dom, 200)
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    dargs = [to_dev(F[n], device) for n in NAMES]
    for tile in (0, 1, 2):
        for hs in SCHEMES:
            for vs in SCHEMES:
                for tag in (gt.H_TAG, gt.W_TAG, gt.V_TAG):
                    for AH in (-1.0, 3.0):
                        outs = (hs + vs + tile) % 2 == 0
                        f = F["f"].copy()
                        hires = np.zeros_like(f) if outs else None
                        advres = np.zeros_like(f) if outs else None
                        rc, _ = orc.do_advection_3d(F["dt"], f, *(F[n] for n in NAMES),
                                                    gt.HALFSPLIT, hs, vs, AH, tag, 1, hires,
                                                    advres)
                        assert rc == 0
                        ctx.set_tuning(fused=1, tile=tile)
                        tf = to_dev(F["f"], device)
                        th = to_dev(np.zeros_like(f), device) if outs else None
                        ta = to_dev(np.zeros_like(f), device) if outs else None
                        ctx.do_advection_3d(F["dt"], tf, *dargs, gt.HALFSPLIT, hs, vs, AH, tag,
                                            th, ta)
                        got = to_host(tf)
                        e = rel_err(got, f)
                        assert e <= TOL, (tile, hs, vs, tag, AH, e)
                        assert np.array_equal(got, f), (tile, hs, vs, tag, AH, "not bit-exact")
                        if outs:
                            assert np.array_equal(to_host(th), hires), (tile, hs, vs, tag)
                            assert rel_err(to_host(ta), advres) <= TOL, (tile, hs, vs, tag)
    ctx.set_tuning(fused=1, tile=0)


def test_fused_f64_bit_exact_wide(device):
    """several strips across i and j (NX = 154: strips at i = 0, 60, 120;
    NY = 134: five to seventeen row strips)"""
    dom = random_domain(150, 130, 6, 7, "islands")
    F = random_fields(dom, 200)
    ctx = gt.Context.from_domain(dom)
    orc = po.Oracle.from_domain(dom)
    dargs = [to_dev(F[n], device) for n in NAMES]
    # tiles 5 and 9: different strip heights for the two kernels (8/16, 16/8)
    for tile in (0, 1, 2, 5, 9):
        for sch in SCHEMES:
            for tag in (gt.H_TAG, gt.W_TAG):
                for AH in (-1.0, 3.0):
                    outs = (sch + tile) % 2 == 0
                    f = F["f"].copy()
                    hires = np.zeros_like(f) if outs else None
                    advres = np.zeros_like(f) if outs else None
                    rc, _ = orc.do_advection_3d(F["dt"], f, *(F[n] for n in NAMES),
                                                gt.HALFSPLIT, sch, sch, AH, tag, 1, hires, advres)
                    assert rc == 0
                    ctx.set_tuning(fused=1, tile=tile)
                    tf = to_dev(F["f"], device)
                    th = to_dev(np.zeros_like(f), device) if outs else None
                    ta = to_dev(np.zeros_like(f), device) if outs else None
                    ctx.do_advection_3d(F["dt"], tf, *dargs, gt.HALFSPLIT, sch, sch, AH, tag,
                                        th, ta)
                    got = to_host(tf)
                    assert np.array_equal(got, f), (tile, sch, tag, AH, rel_err(got, f))
                    if outs:
                        assert np.array_equal(to_host(th), hires), (tile, sch, tag)
                        assert rel_err(to_host(ta), advres) <= TOL, (tile, sch, tag)
    ctx.set_tuning(fused=1, tile=0)


def test_batch_f64_two_tracers_bit_exact(device):
    tr = inputs.ts_case(seed=4, imax=125, jmax=38, kmax=9)
    d = tr.dom
    ctx = gt.Context.from_domain(d)
    orc = po.Oracle.from_domain(d)
    T, S = tr.tracers["T"].copy(), tr.tracers["S"].copy()
    host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn]
    dv = dict(zip(NAMES, [to_dev(a, device) for a in host]))
    fin = to_dev(np.stack([T, S]), device)
    fout = torch.empty_like(fin)
    for _ in range(3):
        for f in (T, S):
            orc.do_advection_3d(tr.dt, f, *host, gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM, -1.0,
                                gt.H_TAG)
        gt.adv3d_batch([dict(ctx=ctx, f_in=fin, f_out=fout, **dv)], 2, tr.dt, gt.P2_PDM,
                       gt.P2_PDM)
        fin, fout = fout, fin
    got = to_host(fin)
    assert np.array_equal(got[0], T) and np.array_equal(got[1], S)


def normwise(a, b):
    scale = np.max(np.abs(b))
    d = np.abs(a - b)
    return float(d.max() / scale), float(np.sqrt(np.mean(d ** 2)) / scale)


@pytest.mark.parametrize("rows,group,npassive", [
    ((16, 16), 1, 4), ((32, 32), 3, 4), ((8, 8), 6, 4),
    ((8, 16), 9, 7),      # the bench's tuning: 9 tracers per wavefront, 8/16 rows
    ((8, 32), 9, 9),      # 11 tracers: a full group of 9 and a partial one of 2
    ((16, 8), 3, 5),      # 7 tracers: groups of 3, 3 and 1
    ((32, 16), 6, 6)])    # 8 tracers: groups of 6 and 2
def test_batch_f32_multidomain(device, rows, group, npassive):
    """T, S and passive tracers on uneven subdomains in one launch, FP32
    against the FP64 oracle over several P2-PDM steps, for every strip-height
    pair and tracer group the tuner can pick, full and partial last groups."""
    tile = gt.tile_for_rows(*rows)
    sizes = [(24, 20), (71, 33), (30, 52)]
    subs = inputs.multitracer_case(seed=9, kmax=8, sizes=sizes, npassive=npassive)
    names = list(subs[0].tracers)
    nt = len(names)
    doms, refs, oracles, tins, touts, ctxs = [], [], [], [], [], []
    for tr in subs:
        d = tr.dom
        ctx = gt.Context.from_domain(d)
        ctxs.append(ctx)
        oracles.append(po.Oracle.from_domain(d))
        stack = np.stack([tr.tracers[n] for n in names])
        refs.append(stack.copy())
        tin = to_dev(stack.astype(np.float32), device)
        tout = torch.empty_like(tin)
        tins.append(tin)
        touts.append(tout)
        host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho]
        doms.append(dict(ctx=ctx, **dict(zip(NAMES[:6], [to_dev(a, device) for a in host]))))
    m0 = [[ctxs[s].budget(to_dev(subs[s].hn, device), tins[s][q]) for q in range(nt)]
          for s in range(len(subs))]
    nsteps = 6
    ctxs[0].set_tuning(fused=1, tile=tile, group=group)
    for step in range(nsteps):
        for s, tr in enumerate(subs):
            host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn]
            for q in range(nt):
                f = refs[s][q]
                oracles[s].do_advection_3d(tr.dt, f, *host, gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM,
                                           -1.0, gt.H_TAG)
        for s in range(len(subs)):
            doms[s]["f_in"], doms[s]["f_out"] = tins[s], touts[s]
        gt.adv3d_batch(doms, nt, subs[0].dt, gt.P2_PDM, gt.P2_PDM)
        tins, touts = touts, tins
    for s, tr in enumerate(subs):
        wet = tr.dom.az == 1
        got = to_host(tins[s]).astype(np.float64)
        for q, n in enumerate(names):
            a = got[q][1:, wet]
            b = refs[s][q][1:, wet]
            emax, erms = normwise(a, b)
            assert emax <= 1e-5, (s, n, emax, erms)
            m1 = ctxs[s].budget(to_dev(tr.hn, device), tins[s][q])[0]
            mass0 = m0[s][q][0]
            assert abs(m1 - mass0) <= 1e-6 * abs(mass0), (s, n, mass0, m1)


@pytest.mark.timeout(600)
def test_batch_f32_multitracer_workload_full_size(device):
    """the batched workload itself: T, S and 16 passive tracers on the eight
    subdomains 64x64 ... 512x384 with 30 levels, at the tuning the bench
    uses (9 tracers per wavefront, 8/16 rows) and at the default, two P2-PDM
    steps against the FP64 oracle: max and RMS error per field normalised
    by the field's largest value <= 1e-5, and each field's FP64 mass budget
    conserved to 1e-6 (the layer heights change: hires != ho)"""
    from oracle import pyoracle as po_
    subs = inputs.multitracer_case(seed=3, kmax=30)
    names = list(subs[0].tracers)
    nt = len(names)
    assert nt == 18 and len(subs) == 8
    po_.set_threads(16)
    refs = [np.stack([tr.tracers[n] for n in names]) for tr in subs]
    nsteps = 2
    for s, tr in enumerate(subs):
        orc = po_.Oracle.from_domain(tr.dom)
        host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn]
        for _ in range(nsteps):
            for q in range(nt):
                orc.do_advection_3d(tr.dt, refs[s][q], *host, gt.HALFSPLIT, gt.P2_PDM,
                                    gt.P2_PDM, -1.0, gt.H_TAG)
    for rows, group in (((8, 16), 9), ((16, 16), 3)):
        ctxs, doms, tins, touts = [], [], [], []
        for tr in subs:
            ctx = gt.Context.from_domain(tr.dom)
            ctxs.append(ctx)
            tin = to_dev(np.stack([tr.tracers[n] for n in names]).astype(np.float32), device)
            tins.append(tin)
            touts.append(torch.empty_like(tin))
            host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho]
            doms.append(dict(ctx=ctx, **dict(zip(NAMES[:6], [to_dev(a, device)
                                                             for a in host]))))
        ctxs[0].set_tuning(fused=1, tile=gt.tile_for_rows(*rows), group=group)
        m0 = [[ctxs[s].budget(to_dev(subs[s].ho, device), tins[s][q])[0] for q in range(nt)]
              for s in range(len(subs))]
        for _ in range(nsteps):
            for s in range(len(subs)):
                doms[s]["f_in"], doms[s]["f_out"] = tins[s], touts[s]
                doms[s]["hires"] = torch.empty_like(doms[s]["ho"])
            gt.adv3d_batch(doms, nt, subs[0].dt, gt.P2_PDM, gt.P2_PDM)
            tins, touts = touts, tins
        worst = (0.0, None)
        for s, tr in enumerate(subs):
            wet = tr.dom.az == 1
            got = to_host(tins[s]).astype(np.float64)
            hires = doms[s]["hires"]
            for q, n in enumerate(names):
                emax, erms = normwise(got[q][1:, wet], refs[s][q][1:, wet])
                print(f"rows {rows} group {group} domain {s} {n}: max {emax:.2e} "
                      f"rms {erms:.2e}")
                assert emax <= 1e-5 and erms <= 1e-6, (rows, group, s, n, emax, erms)
                worst = max(worst, (emax, (s, n)))
                m1 = ctxs[s].budget(hires, tins[s][q])[0]
                assert abs(m1 - m0[s][q]) <= 1e-6 * abs(m0[s][q]), (s, n, m0[s][q], m1)
        print(f"rows {rows} group {group}: worst field {worst}")


@pytest.mark.parametrize("rows,group,npassive,fused", [
    ((8, 16), 9, 7, 1),   # the bench's tuning
    ((16, 32), 3, 5, 1),  # 7 tracers: groups of 3, 3 and 1
    ((8, 8), 6, 4, 1),
    ((16, 16), 3, 3, 0)])  # the per-sub-step kernels: the separate one-pass budget
def test_batch_budgets_formed_in_the_step(device, rows, group, npassive, fused):
    """gt_batch_domain.budget: the FP64 masses of the step's result and its
    volume, formed inside the second strip kernel (lane sums over the strip's
    rows, double-double over lanes and strip-levels), on three uneven
    domains in one launch.  Against the error-free gt_budget_batch_f32 of
    the same result and hires: within 1e-14 relative per tracer and for the
    volume (a lane's FP64 sum of at most 32 positive products and the
    six-level FP64 tree over the wavefront are within 38 roundings of exact,
    4.2e-15; the double-double additions after them lose nothing; all
    tracers here are positive).  The per-sub-step path forms them with gt_budget_batch_f32
    itself: equal bit for bit.  A U-grid tag with a budget is rejected."""
    sizes = [(24, 20), (71, 33), (30, 52)]
    subs = inputs.multitracer_case(seed=11, kmax=7, sizes=sizes, npassive=npassive)
    names = list(subs[0].tracers)
    nt = len(names)
    doms, outs, buds, ctxs = [], [], [], []
    for tr in subs:
        ctx = gt.Context.from_domain(tr.dom)
        ctxs.append(ctx)
        tin = to_dev(np.stack([tr.tracers[n] for n in names]).astype(np.float32), device)
        host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho]
        dd = dict(zip(NAMES[:6], [to_dev(a, device) for a in host]))
        out = torch.empty_like(tin)
        bud = torch.full((nt + 1,), float("nan"), dtype=torch.float64, device=device)
        doms.append(dict(ctx=ctx, f_in=tin, f_out=out, hires=torch.empty_like(dd["ho"]),
                         budget=bud, **dd))
        outs.append(out)
        buds.append(bud)
    ctxs[0].set_tuning(fused=fused, tile=gt.tile_for_rows(*rows), group=group)
    gt.adv3d_batch(doms, nt, subs[0].dt, gt.P2_PDM, gt.P2_PDM)
    torch.cuda.synchronize()
    for s in range(len(subs)):
        exact = ctxs[s].budget_batch(doms[s]["hires"], outs[s]).cpu().numpy()
        got = buds[s].cpu().numpy()
        if not fused:
            assert np.array_equal(got, exact), s
            continue
        err = np.abs(got - exact) / np.abs(exact)
        assert err.max() <= 1e-14, (s, err.max())
        assert np.all(np.isfinite(got))
    with pytest.raises(gt.GtError):
        gt.adv3d_batch(doms, nt, subs[0].dt, gt.P2_PDM, gt.P2_PDM, tag=gt.U_TAG)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_budgets_some_domains(device, dtype):
    """a batch whose first and last domains ask for their budgets and whose
    middle one does not: the w kernel's budget variant runs, the middle
    domain's budget slot is left alone, the budgets asked for equal the
    error-free pass, and every domain's result equals, bit for bit, that of
    the same batch without budgets (the budget-free variant)"""
    sizes = [(24, 20), (71, 33), (30, 52)]
    subs = inputs.multitracer_case(seed=12, kmax=6, sizes=sizes, npassive=1)
    names = list(subs[0].tracers)
    nt = len(names)

    def run(with_budget):
        doms, outs, buds, ctxs = [], [], [], []
        for s, tr in enumerate(subs):
            ctx = gt.Context.from_domain(tr.dom)
            ctxs.append(ctx)
            tin = to_dev(np.stack([tr.tracers[n] for n in names]), device).to(dtype)
            host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho]
            dd = dict(zip(NAMES[:6], [to_dev(a, device) for a in host]))
            out = torch.empty_like(tin)
            bud = torch.full((nt + 1,), 7.0, dtype=torch.float64, device=device)
            ask = with_budget and s != 1
            doms.append(dict(ctx=ctx, f_in=tin, f_out=out, hires=torch.empty_like(dd["ho"]),
                             budget=bud if ask else None, **dd))
            outs.append(out)
            buds.append(bud)
        gt.adv3d_batch(doms, nt, subs[0].dt, gt.P2_PDM, gt.P2_PDM)
        torch.cuda.synchronize()
        return doms, outs, buds, ctxs

    d0, o0, _, _ = run(False)
    d1, o1, b1, c1 = run(True)
    for s in range(len(subs)):
        assert torch.equal(o0[s], o1[s]), s
        assert torch.equal(d0[s]["hires"], d1[s]["hires"]), s
    assert torch.equal(b1[1], torch.full_like(b1[1], 7.0))
    for s in (0, 2):
        if dtype == torch.float32:
            exact = c1[s].budget_batch(d1[s]["hires"], o1[s]).cpu().numpy()
        else:
            exact = np.array([c1[s].budget(d1[s]["hires"], o1[s][q])[0] for q in range(nt)] +
                             [c1[s].budget(d1[s]["hires"], o1[s][0])[1]])
        got = b1[s].cpu().numpy()
        err = np.abs(got - exact) / np.abs(exact)
        assert err.max() <= 1e-13, (s, err.max())


def test_batch_budgets_f64(device):
    """the FP64 batched step (T and S) with its budgets formed in the step,
    against gt_budget of each result (error-free): within 1e-14 relative"""
    tr = inputs.ts_case(seed=4, imax=125, jmax=38, kmax=9)
    d = tr.dom
    ctx = gt.Context.from_domain(d)
    host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho]
    dv = dict(zip(NAMES[:6], [to_dev(a, device) for a in host]))
    fin = to_dev(np.stack([tr.tracers["T"], tr.tracers["S"]]), device)
    fout = torch.empty_like(fin)
    hires = torch.empty_like(dv["ho"])
    bud = torch.zeros(3, dtype=torch.float64, device=device)
    gt.adv3d_batch([dict(ctx=ctx, f_in=fin, f_out=fout, hires=hires, budget=bud, **dv)], 2,
                   tr.dt, gt.P2_PDM, gt.P2_PDM)
    got = bud.cpu().numpy()
    for q in range(2):
        m, v = ctx.budget(hires, fout[q])
        assert abs(got[q] - m) <= 1e-14 * abs(m), (q, got[q], m)
        assert abs(got[2] - v) <= 1e-14 * abs(v)


def test_reserve_covers_batch_budgets(device):
    """after gt_reserve for the tracer counts, the batched steps (FP32 at
    every tracer group, FP64) with their budgets formed in the step allocate
    no device memory: the free memory the device reports is unchanged"""
    tr = inputs.ts_case(seed=4, imax=125, jmax=38, kmax=9)
    d = tr.dom
    ctx = gt.Context.from_domain(d)
    nt32 = 7
    ctx.reserve(nt32, 2)
    host = [tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho]
    dv = dict(zip(NAMES[:6], [to_dev(a, device) for a in host]))
    f64 = to_dev(np.stack([tr.tracers["T"], tr.tracers["S"]]), device)
    o64 = torch.empty_like(f64)
    f32 = torch.stack([f64[0].float()] * nt32).contiguous()
    o32 = torch.empty_like(f32)
    hires = torch.empty_like(dv["ho"])
    b64 = torch.zeros(3, dtype=torch.float64, device=device)
    b32 = torch.zeros(nt32 + 1, dtype=torch.float64, device=device)

    def steps(c):
        for group in (1, 3, 6, 9):
            c.set_tuning(1, 0, group)
            gt.adv3d_batch([dict(ctx=c, f_in=f32, f_out=o32, hires=hires, budget=b32, **dv)],
                           nt32, tr.dt, gt.P2_PDM, gt.P2_PDM)
        gt.adv3d_batch([dict(ctx=c, f_in=f64, f_out=o64, hires=hires, budget=b64, **dv)], 2,
                       tr.dt, gt.P2_PDM, gt.P2_PDM)
        torch.cuda.synchronize()

    # every kernel once on a context of its own first (code loaded on first
    # use takes device memory too), then the reserved context
    warm = gt.Context.from_domain(d)
    steps(warm)
    free0 = torch.cuda.mem_get_info()[0]
    steps(ctx)
    assert torch.cuda.mem_get_info()[0] == free0
    assert b32.abs().min().item() > 0 and b64.abs().min().item() > 0
