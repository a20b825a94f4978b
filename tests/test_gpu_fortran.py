"""The Fortran binding (code_amd/fortran/getm_transport_gpu.F90) end to end:
a Fortran program calls do_advection_3d with the reference's argument list
on host arrays; the result must match the oracle bit for bit (bound written:
1e-12 relative per cell)."""
import subprocess

import numpy as np
import pytest

from code_amd import build, inputs
from code_amd import transport as gt
from oracle import pyoracle as po
from tests.util import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("split,hs,vs", [(2, 4, 4), (2, 6, 6), (0, 5, 3), (3, 2, 1)])
def test_fortran_driver(tmp_path, device, split, hs, vs):
    exe = build.BUILD / "fortran" / "adv_driver"
    if not exe.exists():
        build.build_fortran()
    tr = inputs.blob_case(imax=34, jmax=30, kmax=7)
    d = tr.dom
    nsteps = 4
    f = tr.tracers["c"].copy()
    hdr = np.array([d.imin, d.imax, d.jmin, d.jmax, d.kmax, nsteps, split, hs, vs, 10],
                   np.int32)
    rhdr = np.array([d.dx, d.dy, tr.dt, -1.0])
    fields = [f, tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn]
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(rhdr.tobytes())
        for m in (d.az, d.au, d.av, d.ax):
            fh.write(np.ascontiguousarray(m, np.int32).tobytes())
        for a in fields:
            fh.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin", np.float64).reshape(2, *d.shape3)
    orc = po.Oracle.from_domain(d)
    hires = np.zeros(d.shape3)
    for _ in range(nsteps):
        rc, _ = orc.do_advection_3d(tr.dt, f, tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho,
                                    tr.hn, split, hs, vs, -1.0, 10, 1, hires, None)
        assert rc == 0
    assert rel_err(out[0], f) <= 1e-12
    assert np.array_equal(out[0], f)
    assert np.array_equal(out[1], hires)


def test_fortran_rows_driver(tmp_path, device):
    """do_eqstate_gpu (Jackett, with alpha/beta), eqstate_internal_pressure_gpu
    (Jackett, both methods, on the rho do_eqstate left), eos_internal_pressure_gpu
    (linear, both methods) and salinity_diffusion_gpu called from Fortran on
    device-resident fields: bit for bit the oracle's."""
    exe = build.BUILD / "fortran" / "rows_driver"
    if not exe.exists():
        build.build_fortran()
    tr = inputs.ts_case(seed=3, imax=31, jmax=27, kmax=9)
    d = tr.dom
    T, S = tr.tracers["T"].copy(), tr.tracers["S"].copy()
    nuh = inputs.uniform(d.shape3, 5, 1e-5, 5e-2)
    dt, cnpar = 600.0, 0.6
    hdr = np.array([d.imin, d.imax, d.jmin, d.jmax, d.kmax], np.int32)
    rhdr = np.array([d.dx, d.dy, dt, cnpar])
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(rhdr.tobytes())
        for m in (d.az, d.au, d.av, d.ax):
            fh.write(np.ascontiguousarray(m, np.int32).tobytes())
        for a in (T, S, tr.hn, nuh):
            fh.write(np.ascontiguousarray(a).tobytes())
        sseo = inputs.uniform((d.NY, d.NX), 6, -0.2, 0.2)
        for a in (tr.hun, tr.hvn, d.H, sseo):
            fh.write(np.ascontiguousarray(a).tobytes())
    r = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin", np.float64).reshape(17, *d.shape3)
    orc = po.Oracle.from_domain(d)
    rho, buoy, alpha, beta = (np.zeros(d.shape3) for _ in range(4))
    orc.do_eqstate(3, T, S, tr.hn, rho, buoy, alpha, beta)
    # the pressure methods on Jackett's and on the linear EOS's buoyancy
    ip = []
    lb = np.zeros(d.shape3)
    orc.eos_linear(T, S, np.zeros(d.shape3), lb, inputs.EOS_T0, inputs.EOS_S0,
                   inputs.EOS_DTR0, inputs.EOS_DSR0)
    for b in (buoy, lb):
        o = [np.zeros(d.shape3) for _ in range(6)]
        orc.ip_blumberg_mellor(d.H, tr.hn, tr.hun, tr.hvn, b, *o[:3])
        orc.ip_shchepetkin_mcwilliams(d.H, tr.hn, tr.hun, tr.hvn, b, sseo, *o[3:])
        ip += o
    orc.salinity_diffusion(dt, cnpar, tr.hn, nuh, S)
    names = ["rho", "buoy", "alpha", "beta", "S"] + [
        f"{e} {n}" for e in ("jackett", "linear")
        for n in ("zz_bm", "idpdx_bm", "idpdy_bm", "zz_sm", "idpdx_sm", "idpdy_sm")]
    for q, (name, want) in enumerate(zip(names, [rho, buoy, alpha, beta, S] + ip)):
        assert np.array_equal(out[q], want), (name, rel_err(out[q], want))
    assert np.abs(ip[7]).max() > 0 and np.abs(ip[1]).max() > 0


@pytest.mark.parametrize("scheme", [1, 4, 6])
def test_fortran_external_routines(tmp_path, device, scheme):
    """The external routines with the reference's own argument lists
    (code_amd/fortran/getm_external.F90), called from Fortran as
    do_advection and do_internal_pressure call them: adv_split_u with the H
    grid's and the U grid's (shifted, NX-1 columns) flux masks, adv_split_v
    with the NY-1-row view of the flux mask, adv_split_w, adv_upstream_2dh
    and the argument-less ip_blumberg_mellor() / ip_shchepetkin_mcwilliams()
    on the host model's module variables: bit for bit the oracle's routines
    with the masks of init_advection."""
    exe = build.BUILD / "fortran" / "ext_driver"
    if not exe.exists():
        build.build_fortran()
    tr = inputs.ts_case(seed=5, imax=29, jmax=23, kmax=6)
    d = tr.dom
    k = 3
    dt, AH, splitfac = tr.dt, 2.0, 0.5
    orc = po.Oracle.from_domain(d)
    f = tr.tracers["T"].copy()
    rho, buoy = np.zeros(d.shape3), np.zeros(d.shape3)
    orc.eos_linear(f, tr.tracers["S"], rho, buoy, inputs.EOS_T0, inputs.EOS_S0,
                   inputs.EOS_DTR0, inputs.EOS_DSR0)
    sseo = inputs.uniform(d.H.shape, 9, -0.3, 0.3)
    hdr = np.array([d.imin, d.imax, d.jmin, d.jmax, d.kmax, scheme, k, 1], np.int32)
    rhdr = np.array([d.dx, d.dy, dt, AH, splitfac])
    with open(tmp_path / "in.bin", "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(rhdr.tobytes())
        for m in (d.az, d.au, d.av, d.ax):
            fh.write(np.ascontiguousarray(m, np.int32).tobytes())
        for a in (d.H, sseo, f, tr.uu, tr.vv, tr.ww, tr.hun, tr.hvn, tr.ho, tr.hn, buoy):
            fh.write(np.ascontiguousarray(a, np.float64).tobytes())
    r = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(tmp_path / "out.bin", np.float64)
    n2, n3 = int(np.prod(d.H.shape)), int(np.prod(d.shape3))
    sizes = [3 * n2] * 4 + [3 * n3] * 3
    assert raw.size == sum(sizes)
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    two = [p.reshape(3, *d.H.shape) for p in parts[:4]]
    three = [p.reshape(3, *d.shape3) for p in parts[4:]]

    def slice_case(call, *args, tag=gt.H_TAG):
        fi, Di, adv = f[k].copy(), tr.ho[k].copy(), np.zeros(d.H.shape)
        call(dt, np.ascontiguousarray(f[k]), fi, Di, adv, *args, tag=tag)
        return fi, Di, adv

    want = [
        slice_case(lambda *a, tag: orc.adv_split_u(*a[:5], np.ascontiguousarray(tr.uu[k]),
                                                    np.ascontiguousarray(tr.hun[k]), splitfac,
                                                    scheme, AH, tag)),
        slice_case(lambda *a, tag: orc.adv_split_v(*a[:5], np.ascontiguousarray(tr.vv[k]),
                                                    np.ascontiguousarray(tr.hvn[k]), splitfac,
                                                    scheme, AH, tag)),
        slice_case(lambda *a, tag: orc.adv_split_u(*a[:5], np.ascontiguousarray(tr.uu[k]),
                                                    np.ascontiguousarray(tr.hun[k]), splitfac,
                                                    scheme, AH, tag), tag=gt.U_TAG),
        slice_case(lambda *a, tag: orc.adv_upstream_2dh(
            *a[:5], *(np.ascontiguousarray(x[k]) for x in (tr.uu, tr.vv, tr.hn, tr.hun, tr.hvn)),
            AH, tag)),
    ]
    for q, (w, got) in enumerate(zip(want, two)):
        for name, a, b in zip(("fi", "Di", "adv"), w, got):
            assert np.array_equal(a, b), (q, name, rel_err(b, a))
    fi3, hi3, adv3 = f.copy(), tr.ho.copy(), np.zeros(d.shape3)
    orc.adv_split_w(dt, f, fi3, hi3, adv3, tr.ww, 1.0, scheme, gt.H_TAG, 1)
    for name, a, b in zip(("fi", "hi", "adv"), (fi3, hi3, adv3), three[0]):
        assert np.array_equal(a, b), ("w", name, rel_err(b, a))
    for q, fn in enumerate((orc.ip_blumberg_mellor, orc.ip_shchepetkin_mcwilliams)):
        zz, ix, iy = (np.zeros(d.shape3) for _ in range(3))
        args = (d.H, tr.hn, tr.hun, tr.hvn, buoy) + ((sseo,) if q else ())
        fn(*args, zz, ix, iy)
        for name, a, b in zip(("zz", "idpdx", "idpdy"), (zz, ix, iy), three[1 + q]):
            assert np.array_equal(a, b), ("ip", q, name, rel_err(b, a))
