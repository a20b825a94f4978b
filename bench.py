"""Benchmark of the MI355X transport step.

Workload (one "step"): the FP64 baroclinic transport step of a
400x400x40 GVC domain -- linear EOS, Blumberg-Mellor and
Shchepetkin-McWilliams internal pressure, and P2-PDM HALFSPLIT advection of
temperature and salinity -- with every input resident in HBM.
metric: tracer cell-updates/s (imax*jmax*kmax*ntracers per step / time).

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
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); the transport "
              "step's multi-GPU decomposition is described in DESIGN.md and not run here.",
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
    tr = inputs.config2(imax=a.imax, jmax=a.jmax, kmax=a.kmax)
    d = tr.dom
    st = torch.cuda.Stream(device=dev)
    ctx = gt.Context.from_domain(d, stream=st)
    D = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    uu, vv, ww, hun, hvn, ho, hn = (D(x) for x in (tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho,
                                                   tr.hn))
    H = D(d.H)
    sseo = torch.zeros((d.NY, d.NX), dtype=torch.float64, device=dev)
    T = D(tr.tracers["T"])
    S = D(tr.tracers["S"])
    buoy = torch.zeros_like(T)
    idpdx = torch.zeros_like(T)
    idpdy = torch.zeros_like(T)
    idpdx6 = torch.zeros_like(T)
    idpdy6 = torch.zeros_like(T)
    zz = torch.zeros_like(T)
    zz6 = torch.zeros_like(T)
    hires = torch.zeros_like(T)
    consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
    dt = tr.dt
    torch.cuda.synchronize()

    def eos():
        ctx.eos_linear(T, S, None, buoy, *consts)

    def ip_bm():
        ctx.ip_blumberg_mellor(H, hn, hun, hvn, buoy, zz, idpdx, idpdy)

    def ip_sm():
        ctx.ip_shchepetkin_mcwilliams(H, hn, hun, hvn, buoy, sseo, zz6, idpdx6, idpdy6)

    def adv(f, hr=None):
        ctx.do_advection_3d(dt, f, uu, vv, ww, hun, hvn, ho, hn, gt.HALFSPLIT, gt.P2_PDM,
                            gt.P2_PDM, -1.0, gt.H_TAG, hr, None)

    def step():
        eos()
        ip_bm()
        ip_sm()
        adv(T, hires)
        adv(S)

    with torch.cuda.stream(st):
        for _ in range(a.warmup):
            step()
        st.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step()
        st.synchronize()
        torch.cuda.synchronize()
        t1 = time.perf_counter()

        # component durations with HIP events on the context's stream
        def timed(fn, n):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(n):
                fn()
            e1.record(st)
            e1.synchronize()
            return e0.elapsed_time(e1) * 1e-3 / n

        nrep = max(5, a.steps // 2)
        t_adv = timed(lambda: adv(S), nrep)
        t_bm = timed(ip_bm, nrep)
        t_sm = timed(ip_sm, nrep)
        t_eos = timed(eos, nrep)

    ms = (t1 - t0) / a.steps * 1e3
    ncell = d.imax * d.jmax * d.kmax
    ntr = 2
    value = ncell * ntr / (ms * 1e-3)
    # algorithmic bytes of one tracer's advection step: f in/out + 7 FP64
    # inputs once (uu, vv, ww, hun, hvn, ho, hn)
    b_adv = ncell * (2 * 8 + 7 * 8)
    roof = {
        "bound": "hbm",
        "kernel": "do_advection_3d (one tracer, all sub-steps)",
        "achieved": b_adv / t_adv / 1e9,
        "peak": HBM_PEAK / 1e9,
        "unit": "GB/s",
        "frac": b_adv / t_adv / HBM_PEAK,
        "traffic": None,
    }
    out = {
        "metric": "tracer cell-updates/s (FP64 transport step, T+S)",
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
        "data": "synthetic (seeded GVC basin, streamfunction flow)",
        "config": {"workload": f"{d.imax}x{d.jmax}x{d.kmax} GVC, EOS+BM+SMcW+2-tracer "
                               "P2-PDM HALFSPLIT advection", "imax": d.imax, "jmax": d.jmax,
                   "kmax": d.kmax, "tracers": ntr, "parallelism": "single GPU"},
        "components_ms": {"advection_one_tracer": t_adv * 1e3, "ip_blumberg_mellor": t_bm * 1e3,
                          "ip_shchepetkin_mcwilliams": t_sm * 1e3, "eos_linear": t_eos * 1e3},
        "roofline": roof,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tr, a.cpu_seconds)
    print(json.dumps(out))


def cpu_baseline(tr, seconds):
    """The oracle (C/OpenMP restatement) timed on this host's cores on whole
    steps of the same workload, until `seconds` have elapsed."""
    import numpy as np

    from oracle import pyoracle as po
    from code_amd import inputs
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

    step()  # first touch of the scratch arrays, not timed
    n = 0
    t0 = time.perf_counter()
    while True:
        step()
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    t = (time.perf_counter() - t0) / n
    ncell = d.imax * d.jmax * d.kmax
    return {"value": ncell * 2 / t, "unit": "cell-updates/s", "cores": ncores, "kind": "port",
            "sample": f"{n} whole steps of the same workload (s/step {t:.3f})"}


if __name__ == "__main__":
    main()
