"""Benchmark of the MI355X transport step.

Headline workload (one "step", FP64): the baroclinic transport step of a
400x400x40 GVC basin -- linear EOS, Blumberg-Mellor and
Shchepetkin-McWilliams internal pressure, and P2-PDM HALFSPLIT advection of
temperature and salinity -- with every input resident in HBM.
metric: tracer cell-updates/s = imax*jmax*kmax*2 per step / time.

Second section (batched FP32): T, S and 16 passive tracers on 8 uneven
subdomains (64x64 ... 512x384, 30 layers) in one launch pair per step,
FP32 tracers with FP64 layer heights; its error against the FP64 path after
the timed steps, and its mass-budget drift.

Advection runs as two kernels per step (k_strip_uv: u/2 + v/2, k_strip_wvu:
w + v/2 + u/2).  Their roofline is quoted for the pair, with the algorithmic
bytes of the whole step Ncell*(2*Ntr*s + 7*8): each tracer read and written
once, uu vv ww hu hv ho read and hires written once in FP64.  Strip height
and tracers per wavefront are autotuned (untimed) unless --rows64 (headline)
and --rows32/--group (batched) fix them (rows "R" for both kernels or "R1/R2").

Prints one JSON line.  Run: python bench.py [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "strip_kernels_hbm_traffic.txt")
VALU_FILE = os.path.join(ROOT, "profiles", "strip_kernels_valu_issue.txt")
# wave64 VALU instructions per second: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles
VALU_PEAK = 256 * 4 * 2.4e9 / 4


def parse_rows(v):
    """'R' or 'R1/R2' -> (rows of the first kernel, rows of the second); None if unset"""
    if v in (None, "", "-1"):
        return None
    a, _, b = v.partition("/")
    return int(a), int(b or a)


def rows_str(pair):
    return f"{pair[0]}/{pair[1]}"


def measured_traffic(section, rows, group=None):
    """HBM bytes of one advection step (both strip kernels) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, scaled by what the same
    counters read for a plain copy of known size; None when the summary does
    not cover this strip height / tracer group."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    recs = {}
    for ln in lines:
        if ln.startswith("#") or not ln.strip():
            continue
        head, *kv = ln.split()
        recs.setdefault(head, []).append(dict(x.split("=") for x in kv))
    if "calibration" not in recs or section not in recs:
        return None
    cal = recs["calibration"][0]
    cf = float(cal["copy_read_bytes"]) / (float(cal["fetch_kib"]) * 1024)
    cw = float(cal["copy_write_bytes"]) / (float(cal["write_kib"]) * 1024)
    for r in recs[section]:
        if parse_rows(r["rows"]) != tuple(rows) or (group is not None and
                                                    int(r["group"]) != group):
            continue
        fetch = float(r["uv_fetch_kib"]) + float(r["wvu_fetch_kib"])
        write = float(r["uv_write_kib"]) + float(r["wvu_write_kib"])
        return fetch * 1024 * cf + write * 1024 * cw
    return None
def valu_issue(section, rows, group, seconds):
    """VALU issue rate of one advection step (both strip kernels): the
    wavefront instructions the committed rocprofv3 SQ_INSTS_VALU summary
    counts per dispatch, over this run's HIP-event duration of the pair;
    None when the summary does not cover this strip height / tracer group."""
    try:
        lines = open(VALU_FILE).read().splitlines()
    except OSError:
        return None
    insts = 0.0
    found = set()
    for ln in lines:
        if ln.startswith("#") or not ln.strip():
            continue
        head, *kv = ln.split()
        r = dict(x.split("=") for x in kv)
        if head != section or parse_rows(r["rows"]) != tuple(rows) or int(r["group"]) != group:
            continue
        insts += float(r["valu_insts"])
        found.add(r["kernel"])
    if found != {"uv", "wvu"}:
        return None
    return {"insts_per_step": insts, "achieved": insts / seconds / 1e9,
            "peak": VALU_PEAK / 1e9, "unit": "G wave-instr/s",
            "frac": insts / seconds / VALU_PEAK}


METRIC = "tracer cell-updates/s + % HBM peak; batched FP32 vs plain FP64 at <=1e-5 err"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--imax", type=int, default=400)
    p.add_argument("--jmax", type=int, default=400)
    p.add_argument("--kmax", type=int, default=40)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="bound on the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-adjoint", action="store_true")
    p.add_argument("--no-tidal", action="store_true",
                   help="skip the 2000x1500x50 tidal section (forward and reverse mode)")
    p.add_argument("--no-ext", action="store_true",
                   help="skip the full-EOS and input-producer timings")
    p.add_argument("--serial", action="store_true",
                   help="run EOS + pressure and the advection on one stream")
    p.add_argument("--separate-ip", action="store_true",
                   help="run the EOS and each pressure method as separate launches in the "
                        "step (default: the fused two-launch form)")
    p.add_argument("--shard", action="store_true",
                   help="also time one GPU's share (1000x375x50, 18 FP32 tracers) of the "
                        "2000x1500x50 case")
    p.add_argument("--batched-steps", type=int, default=20)
    p.add_argument("--rows64", default=None,
                   help="FP64 headline: rows per strip, 8, 16 or 32, as R or R1/R2 (default: "
                        "autotuned)")
    p.add_argument("--rows32", default=None,
                   help="FP32 batched: rows per strip as for --rows64 (default: autotuned)")
    p.add_argument("--copy-mb", type=int, default=0,
                   help="also run a plain device copy of this many MiB (counter calibration)")
    p.add_argument("--group", type=int, default=-1,
                   help="FP32 tracers per wavefront 1, 3, 6 or 9 (-1: autotuned)")
    return p.parse_args()


class EventTimer:
    """HIP events around launches on one stream; mean duration in seconds"""

    def __init__(self, torch, stream):
        self.torch, self.st, self.pairs = torch, stream, []

    def __enter__(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.st)
        self.pairs.append([e, None])
        return self

    def __exit__(self, *exc):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record(self.st)
        self.pairs[-1][1] = e

    def mean(self):
        self.st.synchronize()
        ts = [a.elapsed_time(b) * 1e-3 for a, b in self.pairs]
        return sum(ts) / len(ts)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the transport "
              "step is measured on one GPU's subdomain; several GPUs are not run here.",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from code_amd import build as b
    if not (b.BUILD / b.LIB_NAME).exists():
        b.build_hip()
    from code_amd import inputs
    from code_amd import transport as gt

    dev = torch.device("cuda:0")
    if a.copy_mb > 0:
        # a plain copy of known size (read N bytes, write N bytes) whose
        # hardware byte counters calibrate those of the transport kernels
        src = torch.ones(a.copy_mb * 2 ** 18, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    tr = inputs.ts_case(imax=a.imax, jmax=a.jmax, kmax=a.kmax)
    d = tr.dom
    st = torch.cuda.Stream(device=dev)
    ctx = gt.Context.from_domain(d, stream=st)
    # the EOS + internal-pressure chain and the T/S advection read the same T
    # and S and write disjoint fields, so they run on two streams (default
    # priority); --serial puts both on one
    st_ip = st if a.serial else torch.cuda.Stream(device=dev)
    ctx_ip = ctx if a.serial else gt.Context.from_domain(d, stream=st_ip)
    if parse_rows(a.rows64):
        ctx.set_tuning(fused=1, tile=gt.tile_for_rows(*parse_rows(a.rows64)))
    D = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    uu, vv, ww, hun, hvn, ho, hn = (D(x) for x in (tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho,
                                                   tr.hn))
    H = D(d.H)
    sseo = torch.zeros((d.NY, d.NX), dtype=torch.float64, device=dev)
    TS = D(np.stack([tr.tracers["T"], tr.tracers["S"]]))
    TS2 = torch.empty_like(TS)
    shape = TS.shape[1:]
    buoy = torch.zeros(shape, dtype=torch.float64, device=dev)
    idpdx, idpdy, idpdx6, idpdy6, zz, zz6 = (torch.zeros_like(buoy) for _ in range(6))
    hires = torch.zeros_like(buoy)
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    dt = tr.dt
    bufs = [TS, TS2]
    t_adv = EventTimer(torch, st)
    t_ip = EventTimer(torch, st)
    t_ip_other = EventTimer(torch, st)

    tuned = None
    rows64, g64 = parse_rows(a.rows64), 2
    if rows64 is None:
        bt, bg, bms, table = gt.autotune_batch(
            [dict(ctx=ctx, f_in=TS, f_out=TS2, uu=uu, vv=vv, ww=ww, hu=hun, hv=hvn, ho=ho,
                  hires=hires)], 2, dt, gt.P2_PDM, gt.P2_PDM, reps=5, stream=st)
        tuned = {"rows": rows_str(gt.strip_rows(bt)), "tracers_per_wave": bg, "ms": bms,
                 "table_ms": {f"{rows_str(gt.strip_rows(t))}x{g}": v
                              for (t, g), v in table.items()}}
        rows64, g64 = gt.strip_rows(bt), bg

    ev = {}

    def ip_separate(c, s, cur):
        # EOS, then each method from the stored buoyancy (five launches)
        c.eos_linear(cur[0], cur[1], None, buoy, *consts)
        ev["eos"] = torch.cuda.Event()
        ev["eos"].record(s)
        c.ip_blumberg_mellor(H, hn, hun, hvn, buoy, zz, idpdx, idpdy)
        c.ip_shchepetkin_mcwilliams(H, hn, hun, hvn, buoy, sseo, zz6, idpdx6, idpdy6)

    def ip_fused(c, s, cur):
        # both methods with the EOS inside (two launches); T and S are read
        # until the last one ends
        c.eos_internal_pressure(3, cur[0], cur[1], *consts, H, hn, hun, hvn, sseo, zz, idpdx,
                                idpdy, zz6, idpdx6, idpdy6)
        ev["eos"] = torch.cuda.Event()
        ev["eos"].record(s)

    ip_chain = ip_separate if a.separate_ip else ip_fused

    def advect(cur, nxt):
        gt.adv3d_batch([dict(ctx=ctx, f_in=cur, f_out=nxt, uu=uu, vv=vv, ww=ww, hu=hun,
                             hv=hvn, ho=ho, hires=hires)], 2, dt, gt.P2_PDM, gt.P2_PDM,
                       stream=st)

    def step():
        cur, nxt = bufs
        if not a.serial:
            # IP reads cur, which the previous advection wrote; this
            # advection overwrites nxt, which the previous EOS read
            if "adv" in ev:
                st_ip.wait_event(ev["adv"])
            if "eos" in ev:
                st.wait_event(ev["eos"])
        ip_chain(ctx_ip, st_ip, cur)
        advect(cur, nxt)
        ev["adv"] = torch.cuda.Event()
        ev["adv"].record(st)
        bufs.reverse()

    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        for _ in range(a.warmup):
            step()
        st.synchronize()
        st_ip.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        st.synchronize()
        st_ip.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # the components alone, one stream, for the roofline and the split:
        # HIP events on the stream the kernels run on
        for _ in range(a.steps):
            cur, nxt = bufs
            with t_adv:
                advect(cur, nxt)
            with t_ip:
                ip_chain(ctx, st, cur)
            with t_ip_other:
                (ip_fused if a.separate_ip else ip_separate)(ctx, st, cur)
            bufs.reverse()
        st.synchronize()
    ms = (t1 - t0) / a.steps * 1e3
    ncell = d.imax * d.jmax * d.kmax
    ntr = 2
    value = ncell * ntr / (ms * 1e-3)
    adv_s = t_adv.mean()
    ip_s = t_ip.mean()
    # algorithmic bytes of one advection step (T and S, FP64): each tracer
    # read and written once, uu vv ww hu hv ho read and hires written once
    b_adv = ncell * (ntr * 2 * 8 + 7 * 8)
    roof = {"bound": "hbm", "kernel": f"k_strip_uv + k_strip_wvu <double, P2-PDM, {g64} "
                                      "tracer(s) per wave> (one advection step, two launches)",
            "achieved": b_adv / adv_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": b_adv / adv_s / HBM_PEAK, "traffic": measured_traffic("headline", rows64, g64),
            "valu_issue": valu_issue("headline", rows64, g64, adv_s),
            "bytes_per_launch": b_adv, "launch_us": adv_s * 1e6,
            "bytes_per_cell_tracer": b_adv / (ncell * ntr)}
    # EOS + both internal-pressure methods, algorithmic bytes: T S hn hun hvn
    # read, zz idpdx idpdy of each method written, 88 B/cell (the separate
    # chain also writes and reads back the buoyancy, 16 B/cell more)
    ip_o = t_ip_other.mean()
    ip_fused_s, ip_sep_s = (ip_o, ip_s) if a.separate_ip else (ip_s, ip_o)
    b_ip = ncell * 88
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded GVC basin, streamfunction flow, T/S profiles)",
        "config": {"workload": f"{d.imax}x{d.jmax}x{d.kmax} GVC basin: linear EOS + "
                               "Blumberg-Mellor + Shchepetkin-McWilliams + P2-PDM HALFSPLIT "
                               "advection of T and S", "imax": d.imax, "jmax": d.jmax,
                   "kmax": d.kmax, "tracers": ntr, "parallelism": "single GPU",
                   "wet_cell_fraction": float((d.az[2:-2, 2:-2] == 1).mean())},
        "fp64_advection_cell_updates_per_s": ncell * ntr / adv_s,
        "components_ms": {"advection_T_S": adv_s * 1e3, "eos_and_pressure": ip_s * 1e3,
                          "eos_and_pressure_fused": ip_fused_s * 1e3,
                          "eos_and_pressure_separate": ip_sep_s * 1e3,
                          "eos_and_pressure_in_step": "separate" if a.separate_ip else "fused",
                          "timing": "each component alone on one stream, after the timed "
                                    "steps"},
        "streams": 1 if a.serial else 2,
        "eos_and_pressure_GBps": b_ip / ip_fused_s / 1e9,
        "tuning": tuned,
        "roofline": roof,
    }
    del TS2, idpdx6, idpdy6
    progress(f"headline done: {ms:.4f} ms/step")
    if not a.no_batched:
        out["batched_fp32"] = batched(torch, np, gt, inputs, dev, a, ncell * ntr / adv_s)
        progress("batched section done")
    if a.shard:
        out["shard"] = shard_section(torch, np, gt, inputs, dev, a.batched_steps)
        progress("shard section done")
    if not a.no_tidal:
        out["tidal_2000x1500x50"] = tidal_full_section(torch, np, gt, inputs, dev, 5)
        progress("tidal 2000x1500x50 section done")
    if not a.no_ext:
        out["extensions"] = extensions_section(torch, np, gt, inputs, dev, tr, ctx, st,
                                               a.steps)
        progress("extensions done")
    if not a.no_adjoint:
        out["adjoint"] = adjoint_section(torch, np, gt, inputs, dev)
        progress("adjoint done")
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tr, a.cpu_seconds)
        progress("cpu baseline done")
    print(json.dumps(out))


def progress(msg):
    """one line per finished section on stderr (the JSON line is on stdout)"""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


class _Null:
    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def batched(torch, np, gt, inputs, dev, a, fp64_rate):
    """18 tracers on 8 uneven subdomains, one fused FP32 launch per step"""
    subs = inputs.multitracer_case(npassive=16)
    names = list(subs[0].tracers)
    nt = len(names)
    st = torch.cuda.Stream(device=dev)
    doms, doms64, ctxs, fin32, fin64 = [], [], [], [], []
    ncell = 0
    for tr in subs:
        d = tr.dom
        ncell += d.imax * d.jmax * d.kmax
        ctx = gt.Context.from_domain(d, stream=st)
        if parse_rows(a.rows32):
            ctx.set_tuning(1, gt.tile_for_rows(*parse_rows(a.rows32)),
                           a.group if a.group > 0 else 3)
        ctxs.append(ctx)
        stack = np.stack([tr.tracers[n] for n in names])
        f64 = torch.from_numpy(stack).to(dev)
        f32 = f64.to(torch.float32)
        fin64.append([f64, torch.empty_like(f64)])
        fin32.append([f32, torch.empty_like(f32)])
        vel = {k: torch.from_numpy(np.ascontiguousarray(getattr(tr, k2))).to(dev)
               for k, k2 in (("uu", "uu"), ("vv", "vv"), ("ww", "ww"), ("hu", "hun"),
                             ("hv", "hvn"), ("ho", "ho"))}
        vel["hn"] = torch.from_numpy(tr.hn).to(dev)
        vel["hires"] = torch.zeros_like(vel["hn"])
        doms.append(dict(ctx=ctx, **vel))
        tr.tracers.clear()
    dt = subs[0].dt
    hs = vs = gt.P2_PDM
    # FP64 budgets of the initial state (for the drift check)
    m0 = [[ctxs[s].budget(doms[s]["hn"], fin64[s][0][q])[0] for q in range(nt)]
          for s in range(len(subs))]

    # the FP64 budgets (masses of all tracers and the volume) of each
    # domain's result, formed inside the step when asked for
    budgets = [torch.zeros(nt + 1, dtype=torch.float64, device=dev) for _ in doms]

    def launch(bufs, dtype_doms, with_budget=False):
        gt.adv3d_batch([dict(ctx=dd["ctx"], f_in=bufs[s][0], f_out=bufs[s][1], uu=dd["uu"],
                             vv=dd["vv"], ww=dd["ww"], hu=dd["hu"], hv=dd["hv"], ho=dd["ho"],
                             hires=dd["hires"], budget=budgets[s] if with_budget else None)
                        for s, dd in enumerate(dtype_doms)], nt, dt, hs, vs, stream=st)
        for bb in bufs:
            bb.reverse()

    tuned = None
    rows32, group = parse_rows(a.rows32), a.group
    if rows32 is None or a.group < 0:
        bt, bg, bms, table = gt.autotune_batch(
            [dict(ctx=dd["ctx"], f_in=fin32[s][0], f_out=fin32[s][1], uu=dd["uu"], vv=dd["vv"],
                  ww=dd["ww"], hu=dd["hu"], hv=dd["hv"], ho=dd["ho"], hires=dd["hires"])
             for s, dd in enumerate(doms)], nt, dt, hs, vs, reps=3, stream=st)
        tuned = {"rows": rows_str(gt.strip_rows(bt)), "tracers_per_wave": bg, "ms": bms,
                 "table_ms": {f"{rows_str(gt.strip_rows(t))}x{g}": v
                              for (t, g), v in table.items()}}
        rows32, group = gt.strip_rows(bt), bg
    timer, timer_b = EventTimer(torch, st), EventTimer(torch, st)
    nsteps = a.batched_steps
    with torch.cuda.stream(st):
        for _ in range(2):   # warm-up: two steps forward and back is not neutral,
            launch(fin32, doms)  # so the FP64 reference takes the same two steps
            launch(fin64, doms)
        st.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            with timer:
                launch(fin32, doms)
        st.synchronize()
        t1 = time.perf_counter()
        # the same steps with the budgets of every tracer formed in the step
        mb = []
        for _ in range(nsteps):
            with timer_b:
                launch(fin32, doms, with_budget=True)
            mb.append(torch.stack(budgets).clone())
        st.synchronize()
        st.synchronize()
    # the plain FP64 step on the same workload (the reference-exact kernels,
    # tuned for themselves: the FP32 timings are done), timed over the steps
    # that bring the FP64 reference level with the FP32 result
    bt64, bg64, bms64, _ = gt.autotune_batch(
        [dict(ctx=dd["ctx"], f_in=fin64[s][0], f_out=fin64[s][1], uu=dd["uu"], vv=dd["vv"],
              ww=dd["ww"], hu=dd["hu"], hv=dd["hv"], ho=dd["ho"], hires=dd["hires"])
         for s, dd in enumerate(doms)], nt, dt, hs, vs, reps=3, stream=st)
    timer64 = EventTimer(torch, st)
    with torch.cuda.stream(st):
        for _ in range(2 * nsteps):
            with timer64:
                launch(fin64, doms)
        st.synchronize()
    kern = timer.mean()
    kern64 = timer64.mean()
    kern_b = timer_b.mean()
    wall = (t1 - t0) / nsteps
    mb = torch.stack(mb).cpu().numpy()  # (steps, domains, nt + 1)
    drift_b = float(np.max(np.abs(mb[-1, :, :nt] - mb[0, :, :nt]) / np.abs(mb[0, :, :nt])))
    emax, erms, drift = 0.0, 0.0, 0.0
    per_field = {n: [0.0, 0.0] for n in names}
    for s in range(len(subs)):
        wet = torch.from_numpy(subs[s].dom.az == 1).to(dev)
        g = fin32[s][0].to(torch.float64)
        r = fin64[s][0]
        for q in range(nt):
            gq = g[q, 1:][:, wet]
            rq = r[q, 1:][:, wet]
            scale = float(rq.abs().max())
            dq = (gq - rq).abs()
            e1, e2 = float(dq.max()) / scale, float(torch.sqrt(torch.mean(dq * dq))) / scale
            emax, erms = max(emax, e1), max(erms, e2)
            pf = per_field[names[q]]
            pf[0], pf[1] = max(pf[0], e1), max(pf[1], e2)
            m1 = ctxs[s].budget(doms[s]["hn"], fin32[s][0][q])[0]
            drift = max(drift, abs(m1 - m0[s][q]) / abs(m0[s][q]))
    # algorithmic bytes: each FP32 tracer read and written, 6 FP64 fields read
    # and hires written once
    bytes_step = ncell * (nt * 2 * 4 + 7 * 8)
    rate = ncell * nt / kern
    return {
        "workload": "T, S + 16 passive tracers on 8 subdomains (64x64 ... 512x384) x 30 "
                    "layers, P2-PDM HALFSPLIT, FP32 tracers / FP64 layer heights, one launch "
                    "pair per step",
        "cells": ncell, "tracers": nt,
        "value": rate, "unit": "cell-updates/s", "ms_per_step": kern * 1e3,
        "wall_ms_per_step": wall * 1e3,
        "with_budgets": {"ms_per_step": kern_b * 1e3, "cell_updates_per_s": ncell * nt / kern_b,
                         "mass_drift_max_rel": drift_b,
                         "note": "the FP64 mass of every tracer and the volume of every "
                                 "domain formed inside the step (gt_batch_domain.budget), "
                                 f"{nsteps} steps; drift between the first and the last"},
        "tuning": tuned,
        "roofline": {"bound": "hbm", "kernel": "k_strip_uv + k_strip_wvu <float, P2-PDM>",
                     "achieved": bytes_step / kern / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": bytes_step / kern / HBM_PEAK,
                     "traffic": measured_traffic("batched", rows32, group),
                     "valu_issue": valu_issue("batched", rows32, group, kern),
                     "bytes_per_launch": bytes_step},
        "vs_fp64_advection": rate / fp64_rate,
        "fp64_same_workload": {"ms_per_step": kern64 * 1e3,
                               "cell_updates_per_s": ncell * nt / kern64,
                               "rows": rows_str(gt.strip_rows(bt64)),
                               "tracers_per_wave": bg64,
                               "note": "the reference-exact FP64 kernels on these 18 tracers "
                                       "and 8 subdomains, autotuned"},
        "vs_fp64_same_workload": kern64 / kern,
        "err_vs_fp64": {"max": emax, "rms": erms, "steps": 2 * nsteps + 2,
                        "norm": "|f32-f64| / max|f64| over wet cells, worst tracer",
                        "per_field": {n: {"max": v[0], "rms": v[1]}
                                      for n, v in per_field.items()}},
        "mass_drift_max_rel": drift,
    }


def extensions_section(torch, np, gt, inputs, dev, tr, ctx, st, steps):
    """The callers either side of the step on the headline domain, each alone
    on one stream with HIP events: the full equation of state (Jackett et al.
    2006 with the interface expansion coefficients) and the input producers
    (start_macro, general coordinates, ww_momentum_3d), and the vertical
    diffusion after the advection (salinity; temperature with short-wave).
    Algorithmic bytes per cell over the whole (halo-included) arrays: EOS
    read T S hn, write rho buoy alpha beta = 56 B; coordinates read hn hun
    hvn gga, write ho hn huo hun hvo hvn zc = 88 B; ww read uu vv ho hn,
    write ww = 40 B; salinity diffusion read S hn nuh, write S = 32 B;
    temperature + rad = 40 B."""
    d = tr.dom
    n3 = d.NX * d.NY * d.kmax
    D = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    T, S, hn = D(tr.tracers["T"]), D(tr.tracers["S"]), D(tr.hn)
    rho, buoy, alpha, beta = (torch.zeros_like(T) for _ in range(4))
    t_eos = EventTimer(torch, st)
    H, HU, HV = D(d.H), D(d.HU), D(d.HV)
    sh2 = (d.NY, d.NX)
    e2 = {n: torch.zeros(sh2, dtype=torch.float64, device=dev)
          for n in ("sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn", "Dn", "Dun", "Dvn", "Uint",
                    "Vint")}
    h3 = {n: torch.zeros_like(T) for n in ("ho", "hn", "huo", "hun", "hvo", "hvn", "zc")}
    z = D(0.2 * inputs.uniform(sh2, 9, -1, 1) * (d.az > 0))
    uu, vv, ww = D(tr.uu), D(tr.vv), torch.zeros_like(T)
    ctx.set_coordinates(3, ddu=1.0, ddl=1.0, d_gamma=20.0, gamma_surf=True, min_depth=0.5)
    t_sm, t_co, t_ww = (EventTimer(torch, st) for _ in range(3))
    names2 = ("sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn", "Dn", "Dun", "Dvn", "Uint", "Vint")
    names_e = ("sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn")
    names_h = ("ho", "hn", "huo", "hun", "hvo", "hvn", "zc")

    def producers(timed):
        for tm, fn in ((t_sm, lambda: ctx.start_macro(30, 0.5, H, HU, HV, z,
                                                       *(e2[n] for n in names2))),
                       (t_co, lambda: ctx.coordinates(tr.dt, 3600.0, 50.0, H, HU, HV,
                                                      *(e2[n] for n in names_e),
                                                      *(h3[n] for n in names_h))),
                       (t_ww, lambda: ctx.ww_momentum_3d(tr.dt, uu, vv, h3["ho"], h3["hn"],
                                                         ww))):
            with (tm if timed else _Null()):
                fn()

    nuh = D(inputs.uniform(T.shape, 3, 1e-5, 5e-2))
    sh2t = (d.NY, d.NX)
    swr = D(inputs.uniform(sh2t, 5, 0.0, 600.0))
    shf = D(inputs.uniform(sh2t, 6, -200.0, 100.0))
    Aa, g1, g2 = (D(np.full(sh2t, v)) for v in (0.58, 0.35, 23.0))  # Jerlov type I
    dry = D(np.ones(sh2t))
    rad = torch.zeros_like(T)
    Sd, Td = S.clone(), T.clone()
    t_sd, t_td = EventTimer(torch, st), EventTimer(torch, st)

    def diffusion(timed):
        with (t_sd if timed else _Null()):
            ctx.salinity_diffusion(tr.dt, 0.5, hn, nuh, Sd)
        with (t_td if timed else _Null()):
            ctx.temperature_diffusion(tr.dt, 0.5, hn, nuh, swr, shf, Aa, g1, g2, dry, Sd, rad,
                                      Td)

    # momentum advection as a caller of the path: P2-PDM HALFSPLIT on the
    # U and V control volumes (two transport steps and their averaging)
    hun, hvn = D(tr.hun), D(tr.hvn)
    uuEx, vvEx = torch.zeros_like(T), torch.zeros_like(T)
    t_uv = EventTimer(torch, st)

    def momentum(timed):
        with (t_uv if timed else _Null()):
            ctx.uv_advect_3d(tr.dt, uu, vv, ww_in, hn, hn, hun, hvn,
                             gt.HALFSPLIT, gt.P2_PDM, gt.P2_PDM, uuEx, vvEx)

    ww_in = D(tr.ww)
    # Jackett et al. feeding both pressure methods: fused (rho read on land
    # only, buoy not stored) and as do_eqstate followed by the two methods
    sseo_j = D(inputs.uniform(sh2, 47, -0.2, 0.2))
    pj = [torch.zeros_like(T) for _ in range(6)]
    t_jf, t_jc = EventTimer(torch, st), EventTimer(torch, st)

    def jackett_pressure(timed):
        with (t_jf if timed else _Null()):
            ctx.eqstate_internal_pressure(3, 3, T, S, rho, H, hn, hun, hvn, sseo_j, *pj)
        with (t_jc if timed else _Null()):
            ctx.do_eqstate(3, T, S, hn, rho, buoy)
            ctx.ip_blumberg_mellor(H, hn, hun, hvn, buoy, *pj[:3])
            ctx.ip_shchepetkin_mcwilliams(H, hn, hun, hvn, buoy, sseo_j, *pj[3:])

    with torch.cuda.stream(st):
        for _ in range(3):
            ctx.do_eqstate(3, T, S, hn, rho, buoy, alpha, beta)
            producers(False)  # the first coordinates call initialises
            diffusion(False)
            momentum(False)
            jackett_pressure(False)
        st.synchronize()
        for _ in range(steps):
            with t_eos:
                ctx.do_eqstate(3, T, S, hn, rho, buoy, alpha, beta)
            producers(True)
            diffusion(True)
            momentum(True)
            jackett_pressure(True)
        st.synchronize()
    # the whole macro step on the device, one stream, in integrate_3d's order
    # (m3d.F90:446-607): start_macro, coordinates, ww_momentum_3d, T/S
    # advection (strip kernels), salinity and temperature diffusion, linear
    # EOS fused into Blumberg-Mellor and Shchepetkin-McWilliams
    TS = torch.stack([T, S])
    TS2 = torch.empty_like(TS)
    hires = torch.zeros_like(T)
    zz, zz6, ix, iy, ix6, iy6 = (torch.zeros_like(T) for _ in range(6))
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    t_full = EventTimer(torch, st)

    def full_step(cur, nxt):
        producers(False)
        gt.adv3d_batch([dict(ctx=ctx, f_in=cur, f_out=nxt, uu=uu, vv=vv, ww=ww, hu=h3["hun"],
                             hv=h3["hvn"], ho=h3["ho"], hires=hires)], 2, tr.dt, gt.P2_PDM,
                       gt.P2_PDM, stream=st)
        ctx.salinity_diffusion(tr.dt, 0.5, h3["hn"], nuh, nxt[1])
        ctx.temperature_diffusion(tr.dt, 0.5, h3["hn"], nuh, swr, shf, Aa, g1, g2, dry, nxt[1],
                                  rad, nxt[0])
        ctx.eos_internal_pressure(3, nxt[0], nxt[1], *consts, H, h3["hn"], h3["hun"],
                                  h3["hvn"], e2["sseo"], zz, ix, iy, zz6, ix6, iy6)

    bufs = [TS, TS2]
    with torch.cuda.stream(st):
        for _ in range(2):
            full_step(*bufs)
            bufs.reverse()
        for _ in range(steps):
            with t_full:
                full_step(*bufs)
            bufs.reverse()
        st.synchronize()
    e_s, c_s, w_s, m_s = t_eos.mean(), t_co.mean(), t_ww.mean(), t_sm.mean()
    sd_s, td_s, uv_s = t_sd.mean(), t_td.mean(), t_uv.mean()
    full_s = t_full.mean()
    jf_s, jc_s = t_jf.mean(), t_jc.mean()
    gb = lambda b, s: b / s / 1e9  # noqa: E731
    return {
        "workload": f"{d.imax}x{d.jmax}x{d.kmax} (arrays {d.NX}x{d.NY}x{d.kmax + 1})",
        "eqstate_jackett": {"kernel": "k_eqstate_column<3> (rho, buoy, alpha, beta)",
                            "ms": e_s * 1e3, "bytes_per_cell": 56, "GBps": gb(56 * n3, e_s),
                            "frac": 56 * n3 / e_s / HBM_PEAK},
        "coordinates_general": {"kernel": "k_gvc_step (ho hn huo hun hvo hvn zc)",
                                "ms": c_s * 1e3, "bytes_per_cell": 88,
                                "GBps": gb(88 * n3, c_s), "frac": 88 * n3 / c_s / HBM_PEAK},
        "ww_momentum_3d": {"kernel": "k_ww_momentum", "ms": w_s * 1e3, "bytes_per_cell": 40,
                           "GBps": gb(40 * n3, w_s), "frac": 40 * n3 / w_s / HBM_PEAK},
        "start_macro_ms": m_s * 1e3,
        "salinity_diffusion": {"kernel": "k_vertical_diffusion<false>", "ms": sd_s * 1e3,
                               "bytes_per_cell": 32, "GBps": gb(32 * n3, sd_s),
                               "frac": 32 * n3 / sd_s / HBM_PEAK},
        "uv_advect_3d_ms": uv_s * 1e3,
        "full_step_ms": full_s * 1e3,
        "full_step": "start_macro, general coordinates, ww_momentum_3d, P2-PDM advection of "
                     "T and S, salinity and temperature diffusion, linear EOS fused into "
                     "Blumberg-Mellor and Shchepetkin-McWilliams: one stream, FP64",
        "eqstate_jackett_pressure": {"kernel": "k_ip_columns + k_ip_rows<3,BM,SM> (fused)",
                                     "fused_ms": jf_s * 1e3, "chain_ms": jc_s * 1e3,
                                     "chain": "do_eqstate(3) without alpha/beta, then the two "
                                              "methods' kernels",
                                     "bytes_per_cell": 88, "GBps": gb(88 * n3, jf_s),
                                     "frac": 88 * n3 / jf_s / HBM_PEAK},
        "temperature_diffusion": {"kernel": "k_vertical_diffusion<true> (with short-wave)",
                                  "ms": td_s * 1e3, "bytes_per_cell": 40,
                                  "GBps": gb(40 * n3, td_s),
                                  "frac": 40 * n3 / td_s / HBM_PEAK},
    }


def adjoint_section(torch, np, gt, inputs, dev, nsteps=48, ckpt=8):
    """Reverse mode at 256x256x30 (sigma, Superbee HALFSPLIT): the gradient of
    J = sum_{32x32 box, all k} c^2 (final c) + sum_box w*idpdx(T_final)
    (Blumberg-Mellor of the linear-EOS buoyancy) with respect to the initial c
    and T, by 48 forward steps with a checkpoint every 8 and the reverse sweep;
    velocities and layer heights are prescribed."""
    tr = inputs.adjoint_case()
    d = tr.dom
    st = torch.cuda.Stream(device=dev)
    ctx = gt.Context.from_domain(d, stream=st)
    ctx.set_adjoint_mode(1)  # each segment's sub-step states kept while it is recomputed
    D = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)  # noqa: E731
    uu, vv, ww, hu, hv, ho, hn = (D(x) for x in (tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho,
                                                 tr.hn))
    H = D(d.H)
    c0, T0, S0 = D(tr.tracers["c"]), D(tr.tracers["T"]), D(tr.tracers["S"])
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    j0, i0 = d.NY // 2 - 16, d.NX // 2 - 16
    w = torch.zeros_like(c0)
    w[1:, j0:j0 + 32, i0:i0 + 32] = D(inputs.uniform((d.kmax, 32, 32), 17, -1, 1))
    cF, TF, lam_c, g_c, g_T = (torch.empty_like(c0) for _ in range(5))
    buoy, zz, ix, iy, bbar, Tbar, zero = (torch.zeros_like(c0) for _ in range(7))

    def forward(f0, f):
        f.copy_(f0)
        for _ in range(nsteps):
            ctx.do_advection_3d(tr.dt, f, uu, vv, ww, hu, hv, ho, hn, gt.HALFSPLIT, gt.SUPERBEE,
                                gt.SUPERBEE, -1.0, gt.H_TAG)

    def gradient():
        # c: lam = 2 c_final in the box
        forward(c0, cF)
        lam_c.zero_()
        lam_c[1:, j0:j0 + 32, i0:i0 + 32] = 2.0 * cF[1:, j0:j0 + 32, i0:i0 + 32]
        ctx.adv3d_adjoint(nsteps, ckpt, tr.dt, c0, lam_c, g_c, uu, vv, ww, hu, hv, ho,
                          gt.SUPERBEE, gt.SUPERBEE)
        # T: forward to T_final, IP of its buoyancy, then EOS^T BM^T w and the
        # advection adjoint
        forward(T0, TF)
        ctx.eos_linear(TF, S0, None, buoy, *consts)
        ctx.ip_blumberg_mellor(H, hn, hu, hv, buoy, zz, ix, iy)
        bbar.zero_()
        Tbar.zero_()
        ctx.ip_blumberg_mellor_adjoint(hu, hv, zz, w, zero, bbar)
        ctx.eos_linear_adjoint(bbar, Tbar, None, inputs.EOS_DTR0, inputs.EOS_DSR0)
        ctx.adv3d_adjoint(nsteps, ckpt, tr.dt, T0, Tbar, g_T, uu, vv, ww, hu, hv, ho,
                          gt.SUPERBEE, gt.SUPERBEE)

    # strip heights for one FP64 tracer on this grid (untimed; the context's
    # strip steps, forward and reverse, use them)
    hires = torch.empty_like(c0)
    bt, bg, _, _ = gt.autotune_batch([dict(ctx=ctx, f_in=c0, f_out=cF, uu=uu, vv=vv, ww=ww,
                                           hu=hu, hv=hv, ho=ho, hires=hires)], 1, tr.dt,
                                     gt.SUPERBEE, gt.SUPERBEE, reps=3, stream=st)
    with torch.cuda.stream(st):
        gradient()
        st.synchronize()
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            gradient()
        st.synchronize()
        t = (time.perf_counter() - t0) / reps
    ncell = d.imax * d.jmax * d.kmax
    return {"workload": f"{d.imax}x{d.jmax}x{d.kmax} sigma, Superbee HALFSPLIT, {nsteps} steps, "
                        f"checkpoint every {ckpt}; dJ/dc0 and dJ/dT0 (J: box c^2 + BM term)",
            "ms_per_gradient": t * 1e3, "rows": rows_str(gt.strip_rows(bt)),
            "gradient_cell_steps_per_s": ncell * nsteps * 2 / t,
            "note": "per tracer: 48 forward steps for the cost, then the adjoint call (48 "
                    "forward steps to the checkpoints, the last 8 through the sub-step "
                    "kernels keeping their sub-step states, the other 40 recomputed so from "
                    "the checkpoints, 48 reverse steps of five transposes)",
            "grad_norm_c": float(g_c.norm()), "grad_norm_T": float(g_T.norm())}


def shard_section(torch, np, gt, inputs, dev, steps):
    """One GPU's share of the tidal 2000x1500x50 case split 2x4: a 1000x375x50
    subdomain on general coordinates, an M2 tide (30 micro steps of a
    stand-in 2-D mode per 300 s macro step), the input producers on the device
    (start_macro, coordinates, layer transports, ww_momentum_3d) and T, S and
    16 passive tracers in FP32 advected by the batched step (autotuned);
    HIP events on the stream for the producers and the advection, the FP64
    mass drift of every tracer over the timed steps."""
    from code_amd import tidal
    d = inputs.make_domain(1000, 375, 50, 5, hmin=10.0, hmax=50.0)
    st = torch.cuda.Stream(device=dev)
    ctx = gt.Context.from_domain(d, stream=st)
    nt = 18
    hires = torch.zeros(d.shape3, dtype=torch.float64, device=dev)
    t_prod, t_adv = EventTimer(torch, st), EventTimer(torch, st)
    with torch.cuda.stream(st):
        tide = tidal.TidalShare(d, ctx, dev, M=30, dt=300.0, umax=1.0)
        fin = tidal.initial_tracers(d, dev, nt)
        fout = torch.empty_like(fin)
        dd, hn = tide.macro_step()
        bt, bg, bms, _ = gt.autotune_batch([dict(dd, ctx=ctx, f_in=fin, f_out=fout,
                                                 hires=hires)], nt, tide.dt, gt.P2_PDM,
                                           gt.P2_PDM, reps=3, stream=st)
        bufs = [fin, fout]
        gt.adv3d_batch([dict(dd, ctx=ctx, f_in=bufs[0], f_out=bufs[1], hires=hires)], nt,
                       tide.dt, gt.P2_PDM, gt.P2_PDM, stream=st)
        bufs.reverse()
        m0 = [ctx.budget(hn, bufs[0][q])[0] for q in range(nt)]
        for _ in range(steps):
            tide.micro_steps()  # the stand-in 2-D mode: not timed
            with t_prod:
                dd, hn = tide.producers()
            with t_adv:
                gt.adv3d_batch([dict(dd, ctx=ctx, f_in=bufs[0], f_out=bufs[1], hires=hires)],
                               nt, tide.dt, gt.P2_PDM, gt.P2_PDM, stream=st)
            bufs.reverse()
        st.synchronize()
        m1 = [ctx.budget(hn, bufs[0][q])[0] for q in range(nt)]
        # the FP32 reverse mode over the next 24 macro steps of the tide (the
        # producers' outputs of every step kept on the device), a checkpoint
        # every 6 steps, each segment's layer heights and sub-step states kept
        # while it is recomputed: one untimed call that allocates the work
        # fields, then one timed call
        nadj, ck = 24, 6
        ctx.set_adjoint_mode(1)
        seq = []
        for _ in range(nadj):
            dd_t, _ = tide.macro_step()
            seq.append(tuple(dd_t[n].clone() for n in ("uu", "vv", "ww", "hu", "hv", "ho")))
        f0 = bufs[0]
        gen = torch.Generator(device=dev).manual_seed(7)
        lamT = torch.rand(f0.shape, generator=gen, device=dev, dtype=torch.float32) - 0.5
        lam0, ffin = torch.empty_like(f0), torch.empty_like(f0)
        t_rev = EventTimer(torch, st)
        ctx.adv3d_adjoint_seq(nt, ck, tide.dt, f0, lamT, lam0, seq, gt.P2_PDM, gt.P2_PDM,
                              f_final=ffin)
        with t_rev:
            ctx.adv3d_adjoint_seq(nt, ck, tide.dt, f0, lamT, lam0, seq, gt.P2_PDM, gt.P2_PDM,
                                  f_final=ffin)
        st.synchronize()
        del seq
        # Euler: each step is positively homogeneous of degree 1 in the tracer,
        # so <lam0, f0> = <lamT, f_final>
        lhs = sum(float(torch.dot(lam0[q].reshape(-1).double(), f0[q].reshape(-1).double()))
                  for q in range(nt))
        rhs = sum(float(torch.dot(lamT[q].reshape(-1).double(), ffin[q].reshape(-1).double()))
                  for q in range(nt))
        del lamT, lam0, ffin
    tr_ms = t_rev.mean() * 1e3
    n3 = (d.kmax + 1) * (d.jmax + 4) * (d.imax + 4)
    # heights D1..D5 per step of a segment; tracer sets: 4 sub-step states per
    # step of a segment, lam x2, the checkpoints, the states inside a segment
    work_gb = (5 * n3 * 8 * ck + (4 * ck + 2 + nadj // ck + ck - 1) * nt * n3 * 4) / 1e9
    ta, tp = t_adv.mean(), t_prod.mean()
    ncell = d.imax * d.jmax * d.kmax
    b = ncell * (nt * 2 * 4 + 7 * 8)
    drift = max(abs(b1 - a0) / abs(a0) for a0, b1 in zip(m0, m1) if a0 != 0.0)
    return {"workload": "one GPU's share of the tidal 2000x1500x50 case split 2x4: 1000x375x50 "
                        "general coordinates, M2 tide, producers on the device each 300 s macro "
                        "step, T, S + 16 passive tracers FP32, P2-PDM HALFSPLIT",
            "rows": rows_str(gt.strip_rows(bt)), "tracers_per_wave": bg,
            "ms_per_step": (ta + tp) * 1e3, "advection_ms": ta * 1e3, "producers_ms": tp * 1e3,
            "cell_updates_per_s": ncell * nt / ta, "algorithmic_GBps": b / ta / 1e9,
            "frac_of_8TBps": b / ta / HBM_PEAK, "mass_drift_max_rel": drift,
            "steps": steps,
            "adjoint_f32": {"steps": nadj, "ckpt_every": ck, "tracers": nt, "ms": tr_ms,
                            "ms_per_step": tr_ms / nadj, "work_GB": work_gb,
                            "euler_identity_rel": abs(lhs - rhs) / abs(rhs),
                            "forcing": "24 consecutive macro steps of the M2 tide (the "
                                       "producers' outputs of each step)"}}


def tidal_full_section(torch, np, gt, inputs, dev, steps):
    """The tidal 2000x1500x50 case on one GPU at its stated size: general
    coordinates, an M2 tide (30 micro steps of a stand-in 2-D mode per 300 s
    macro step), the input producers on the device, T, S and 16 passive
    tracers in FP32 advected by the batched step (autotuned) with the FP64
    mass and volume budgets of all 18 tracers formed in the step
    (gt_batch_domain.budget), timed with HIP events on the stream (the
    error-free one-pass gt_budget_batch_f32 of the same result timed beside
    it for comparison); then the
    checkpointed FP32 reverse mode of the 18 tracers over 24 macro steps
    (code_amd/tidal.py TidalAdjointWindow: the tide's state and the tracers
    every 6 steps, a segment's inputs regenerated by the producers in the
    reverse sweep, the library's checkpoint every 2 steps inside a segment, 6
    tracers per call), timed as a whole, with its device memory."""
    import gc

    from code_amd import tidal
    gc.collect()
    torch.cuda.empty_cache()
    d = inputs.make_domain(2000, 1500, 50, 5, hmin=10.0, hmax=50.0)
    st = torch.cuda.Stream(device=dev)
    ctx = gt.Context.from_domain(d, stream=st)
    nt = 18
    hires = torch.zeros(d.shape3, dtype=torch.float64, device=dev)
    t_prod, t_adv, t_bud = (EventTimer(torch, st) for _ in range(3))
    mass = []
    with torch.cuda.stream(st):
        tide = tidal.TidalShare(d, ctx, dev, M=30, dt=300.0, umax=1.0)
        tide.macro_step()  # the coordinates' first call
        fin = tidal.initial_tracers(d, dev, nt)
        snap0 = tide.snapshot()
        f_init = fin.clone()
        fout = torch.empty_like(fin)
        dd, hn = tide.macro_step()
        bt, bg, _, _ = gt.autotune_batch([dict(dd, ctx=ctx, f_in=fin, f_out=fout,
                                               hires=hires)], nt, tide.dt, gt.P2_PDM,
                                         gt.P2_PDM, reps=2, stream=st)
        tide.restore(snap0)
        bufs = [fin, fout]
        for q in range(steps + 1):
            tide.micro_steps()  # the stand-in 2-D mode: not timed
            timed = q > 0
            bud = torch.empty(nt + 1, dtype=torch.float64, device=dev)
            with (t_prod if timed else _Null()):
                dd, hn = tide.producers()
            # the step with the FP64 budgets of all tracers formed in it
            with (t_adv if timed else _Null()):
                gt.adv3d_batch([dict(dd, ctx=ctx, f_in=bufs[0], f_out=bufs[1], hires=hires,
                                     budget=bud)], nt, tide.dt, gt.P2_PDM, gt.P2_PDM, stream=st)
            mass.append(bud)
            # the error-free one-pass budget of the same result, for comparison
            with (t_bud if timed else _Null()):
                exact = ctx.budget_batch(hires, bufs[1])
            bufs.reverse()
        st.synchronize()
        budget_vs_exact = float(((bud - exact).abs() / exact.abs()).max())
        st.synchronize()
        m = torch.stack(mass).cpu().numpy()
        drift = float(np.max(np.abs(m[-1, :nt] - m[0, :nt]) / np.abs(m[0, :nt])))
        del bufs, fin, fout, mass, dd, hn
        # reverse mode over 24 macro steps from f_init
        tide.restore(snap0)
        del snap0
        gc.collect()
        torch.cuda.empty_cache()
        nadj, seg, inner, batch = 24, 6, 2, 6
        win = tidal.TidalAdjointWindow(tide, ctx, nsteps=nadj, seg=seg, inner_ckpt=inner,
                                       batch=batch, stream=st)
        f_fin = torch.empty_like(f_init)
        st.synchronize()
        t0 = time.perf_counter()
        win.forward(f_init, f_fin, hires=hires)
        st.synchronize()
        t1 = time.perf_counter()
        gen = torch.Generator(device=dev).manual_seed(7)
        lam = torch.rand(f_init.shape, generator=gen, device=dev, dtype=torch.float32) - 0.5
        rhs = [float(torch.dot(lam[q].reshape(-1).double(), f_fin[q].reshape(-1).double()))
               for q in range(nt)]
        scale = [float(lam[q].double().norm() * f_fin[q].double().norm()) for q in range(nt)]
        del f_fin
        torch.cuda.empty_cache()
        st.synchronize()
        t2 = time.perf_counter()
        win.reverse(lam)
        st.synchronize()
        t3 = time.perf_counter()
        free, total = torch.cuda.mem_get_info(dev)
        used_gb = (total - free) / 1e9
        lhs = [float(torch.dot(lam[q].reshape(-1).double(), f_init[q].reshape(-1).double()))
               for q in range(nt)]
        euler = max(abs(a - b) / abs(b) for a, b in zip(lhs, rhs))
        euler_cs = max(abs(a - b) / c for a, b, c in zip(lhs, rhs, scale))
        win.release()
        del lam, f_init, win
    ta, tp, tb = t_adv.mean(), t_prod.mean(), t_bud.mean()
    ncell = d.imax * d.jmax * d.kmax
    n3 = (d.kmax + 1) * (d.jmax + 4) * (d.imax + 4)
    b = ncell * (nt * 2 * 4 + 7 * 8)
    set_gb = nt * n3 * 4 / 1e9
    out = {"workload": "tidal 2000x1500x50 general coordinates on one GPU, M2 tide, producers "
                       "on the device each 300 s macro step, T, S + 16 passive tracers FP32, "
                       "P2-PDM HALFSPLIT, FP64 budgets of all tracers each step",
           "rows": rows_str(gt.strip_rows(bt)), "tracers_per_wave": bg, "steps": steps,
           "ms_per_step": (ta + tp) * 1e3, "advection_ms": ta * 1e3,
           "producers_ms": tp * 1e3,
           "budgets": "formed in the step (included in advection_ms)",
           "separate_error_free_budget_ms": tb * 1e3,
           "in_step_budget_vs_error_free_rel": budget_vs_exact,
           "cell_updates_per_s": ncell * nt / ta, "algorithmic_GBps": b / ta / 1e9,
           "frac_of_8TBps": b / ta / HBM_PEAK, "mass_drift_max_rel": drift,
           "tracer_set_GB": set_gb,
           "adjoint_f32": {"steps": nadj, "tide_and_tracer_checkpoint_every": seg,
                           "inner_ckpt_every": inner, "tracers_per_call": batch,
                           "tracers": nt, "forward_ms": (t1 - t0) * 1e3,
                           "reverse_ms": (t3 - t2) * 1e3,
                           "ms_per_step": (t3 - t2 + t1 - t0) * 1e3 / nadj,
                           "device_GB_in_use_after_reverse": used_gb,
                           "euler_identity_rel": euler,
                           "euler_identity_rel_to_norms": euler_cs,
                           "euler_note": "max over tracers of |<lam0,f0> - <lamT,fT>| over "
                                         "|<lamT,fT>| (the random zero-mean lamT makes that "
                                         "dot product small) and over |lamT| |fT|"}}
    ctx.close()
    del ctx, tide, hires
    gc.collect()
    torch.cuda.empty_cache()
    return out


def cpu_baseline(tr, seconds):
    """The oracle (C/OpenMP restatement of the reference) timed on this host's
    cores on whole steps of the same workload, until `seconds` have elapsed."""
    import numpy as np

    from code_amd import inputs
    from oracle import pyoracle as po
    ncores = len(os.sched_getaffinity(0))
    po.set_threads(ncores)
    d = tr.dom
    orc = po.Oracle.from_domain(d)
    T = tr.tracers["T"].copy()
    S = tr.tracers["S"].copy()
    buoy, rho = np.zeros(d.shape3), np.zeros(d.shape3)
    zz, ix, iy = (np.zeros(d.shape3) for _ in range(3))
    zz6, ix6, iy6 = (np.zeros(d.shape3) for _ in range(3))
    sseo = np.zeros((d.NY, d.NX))
    hires = np.zeros(d.shape3)
    args = (tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn)
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)

    def step():
        orc.eos_linear(T, S, rho, buoy, *consts)
        orc.ip_blumberg_mellor(d.H, tr.hn, tr.hun, tr.hvn, buoy, zz, ix, iy)
        orc.ip_shchepetkin_mcwilliams(d.H, tr.hn, tr.hun, tr.hvn, buoy, sseo, zz6, ix6, iy6)
        orc.do_advection_3d(tr.dt, T, *args, 2, 6, 6, -1.0, 10, 1, hires, None)
        orc.do_advection_3d(tr.dt, S, *args, 2, 6, 6, -1.0, 10, 1, None, None)

    def timed(budget):
        n, t0 = 0, time.perf_counter()
        while True:
            step()
            n += 1
            if time.perf_counter() - t0 >= budget:
                return (time.perf_counter() - t0) / n, n

    step()  # first touch of the scratch arrays, not timed
    # three runs of a third of the budget each on all cores, the median kept;
    # then three single steps on one thread
    runs = [timed(seconds / 3.0) for _ in range(3)]
    t = sorted(r[0] for r in runs)[1]
    n = sum(r[1] for r in runs)
    po.set_threads(1)
    t1 = sorted(timed(0.0)[0] for _ in range(3))[1]
    po.set_threads(ncores)
    ncell = d.imax * d.jmax * d.kmax
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(x.split(":", 1)[1].strip() for x in fh if x.startswith("model name"))
    except (OSError, StopIteration):
        pass
    return {"value": ncell * 2 / t, "unit": "cell-updates/s", "cores": ncores, "kind": "port",
            "sample": f"{n} whole steps of the same 400x400x40 workload in three runs, median "
                      f"s/step {t:.4f}, OpenMP threads {ncores}",
            "cpu_model": model, "one_thread_value": ncell * 2 / t1,
            "one_thread_s_per_step": t1}


if __name__ == "__main__":
    main()
