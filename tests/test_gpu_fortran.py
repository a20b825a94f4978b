"""The Fortran binding (code_amd/fortran/getm_transport_gpu.F90) end to end:
a Fortran program calls do_advection_3d with the reference's argument list
on host arrays; the result must match the oracle bit for bit (bound written:
1e-12 relative per cell)."""
import subprocess

import numpy as np
import pytest

from code_amd import build, inputs
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
    """do_eqstate_gpu (Jackett, with alpha/beta) and salinity_diffusion_gpu
    called from Fortran on device-resident fields: bit for bit the oracle's."""
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
    r = subprocess.run([str(exe), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.bin", np.float64).reshape(5, *d.shape3)
    orc = po.Oracle.from_domain(d)
    rho, buoy, alpha, beta = (np.zeros(d.shape3) for _ in range(4))
    orc.do_eqstate(3, T, S, tr.hn, rho, buoy, alpha, beta)
    orc.salinity_diffusion(dt, cnpar, tr.hn, nuh, S)
    for q, (name, want) in enumerate((("rho", rho), ("buoy", buoy), ("alpha", alpha),
                                      ("beta", beta), ("S", S))):
        assert np.array_equal(out[q], want), (name, rel_err(out[q], want))
