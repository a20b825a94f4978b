"""The restart file format (CPU: host entry points of the library).

The files are the reference's NetCDF hotstart layout
(create_restart_ncdf.F90: dimensions xax, yax, zax = kmax+1; the integer
"loop"; FP64 fields over (x, y) and (x, y, z) in Fortran order, i.e.
(yax, xax) and (zax, yax, xax) in the file; the title attribute).  Pinned
against an independent NetCDF-3 implementation (scipy.io.netcdf_file):
what the library writes scipy reads back identically, and what scipy writes
(classic and 64-bit-offset versions) the library reads back identically.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from code_amd import inputs
from code_amd import transport as gt

IMIN, IMAX, JMIN, JMAX, KMAX = 1, 13, 1, 9, 5


def fields(seed=1):
    NX, NY = IMAX - IMIN + 5, JMAX - JMIN + 5
    return {"hn": inputs.uniform((KMAX + 1, NY, NX), seed, 0.5, 2.0),
            "T": inputs.uniform((KMAX + 1, NY, NX), seed + 1, 2.0, 20.0),
            "ssen": inputs.uniform((NY, NX), seed + 2, -0.5, 0.5)}


def interior(a):
    return a[..., 2:-2, 2:-2]


@pytest.mark.parametrize("halos", [False, True])
def test_written_file_reads_back_with_scipy(tmp_path, halos):
    f = fields()
    p = tmp_path / "restart.nc"
    gt.restart_write_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, f, loop=1234, with_halos=halos)
    nc = netcdf_file(str(p), "r", mmap=False)
    try:
        assert nc.version_byte == 1
        assert nc.title == b"GETM NetCDF hotstart file"
        sel = (lambda a: a) if halos else interior
        assert nc.dimensions["xax"] == sel(f["ssen"]).shape[1]
        assert nc.dimensions["yax"] == sel(f["ssen"]).shape[0]
        assert nc.dimensions["zax"] == KMAX + 1
        assert int(nc.variables["loop"].getValue()) == 1234
        assert nc.variables["hn"].dimensions == ("zax", "yax", "xax")
        assert nc.variables["ssen"].dimensions == ("yax", "xax")
        for name, a in f.items():
            assert np.array_equal(nc.variables[name][:], sel(a)), name
    finally:
        nc.close()


@pytest.mark.parametrize("version", [1, 2])
def test_scipy_file_reads_into_the_arrays(tmp_path, version):
    f = fields(7)
    p = tmp_path / "from_scipy.nc"
    nc = netcdf_file(str(p), "w", version=version)
    nc.createDimension("xax", IMAX - IMIN + 1)
    nc.createDimension("yax", JMAX - JMIN + 1)
    nc.createDimension("zax", KMAX + 1)
    nc.createVariable("loop", "i", ()).data[()] = 77
    nc.createVariable("julianday", "i", ()).data[()] = 2451545  # other variables: skipped
    for name, a in f.items():
        dims = ("zax", "yax", "xax") if a.ndim == 3 else ("yax", "xax")
        var = nc.createVariable(name, "d", dims)
        var[:] = interior(a)
        var.units = "m"
    nc.close()
    got = {n: np.full_like(a, -5.0) for n, a in f.items()}
    loop = gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, got)
    assert loop == 77
    for n, a in f.items():
        assert np.array_equal(interior(got[n]), interior(a)), n
        halo = np.ones(a.shape[-2:], bool)
        halo[2:-2, 2:-2] = False
        assert np.all(got[n][..., halo] == -5.0), n  # the halos are kept


def test_round_trip_and_errors(tmp_path):
    f = fields(3)
    p = tmp_path / "rt.nc"
    gt.restart_write_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, f, loop=-3)
    got = {n: np.zeros_like(a) for n, a in f.items()}
    assert gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, got) == -3
    for n in f:
        assert np.array_equal(interior(got[n]), interior(f[n]))
    with pytest.raises(gt.GtError, match="xax"):
        gt.restart_read_host(p, IMIN, IMAX + 1, JMIN, JMAX, KMAX, got)
    with pytest.raises(gt.GtError, match="no FP64 variable S"):
        gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, {"S": got["T"]})
    with pytest.raises(gt.GtError, match="dimensions"):
        gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, {"ssen": got["T"]})
    bad = tmp_path / "bad.nc"
    bad.write_bytes(b"HDF5 not a classic file")
    with pytest.raises(gt.GtError, match="classic"):
        gt.restart_read_host(bad, IMIN, IMAX, JMIN, JMAX, KMAX, got)
