"""The restart file format (CPU: host entry points of the library).

The files are the reference's NetCDF hotstart layout
(create_restart_ncdf.F90: dimensions xax, yax, zax = kmax+1; the scalars
loop, julianday, secondsofday (int) and timestep (double) that
open_restart_ncdf.F90:48-62 requires; FP64 fields over (x, y) and (x, y, z)
in Fortran order, i.e. (yax, xax) and (zax, yax, xax) in the file; the title
attribute).  Pinned
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
    return {"z": inputs.uniform((NY, NX), seed + 3, -0.5, 0.5),
            "hn": inputs.uniform((KMAX + 1, NY, NX), seed, 0.5, 2.0),
            "T": inputs.uniform((KMAX + 1, NY, NX), seed + 1, 2.0, 20.0),
            "ssen": inputs.uniform((NY, NX), seed + 2, -0.5, 0.5)}


def interior(a):
    return a[..., 2:-2, 2:-2]


@pytest.mark.parametrize("halos", [False, True])
def test_written_file_reads_back_with_scipy(tmp_path, halos):
    f = fields()
    p = tmp_path / "restart.nc"
    gt.restart_write_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, f, loop=1234, with_halos=halos,
                          julianday=2460311, secondsofday=43210, timestep=30.5)
    nc = netcdf_file(str(p), "r", mmap=False)
    try:
        assert nc.version_byte == 1
        assert nc.title == b"GETM NetCDF hotstart file"
        # every variable the reference's open_restart_ncdf requires, with the
        # type create_restart_ncdf gives it
        for name, typ, val in (("loop", "i", 1234), ("julianday", "i", 2460311),
                               ("secondsofday", "i", 43210), ("timestep", "d", 30.5)):
            v = nc.variables[name]
            assert v.dimensions == () and v.typecode() == typ, name
            assert v.getValue() == val, name
        assert nc.variables["z"].dimensions == ("yax", "xax")
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
    nc.createVariable("julianday", "i", ()).data[()] = 2451545
    nc.createVariable("secondsofday", "i", ()).data[()] = 3600
    nc.createVariable("timestep", "d", ()).data[()] = 12.25
    nc.createVariable("Uinto", "d", ("yax", "xax"))[:] = 0.0   # other variables: skipped
    for name, a in f.items():
        dims = ("zax", "yax", "xax") if a.ndim == 3 else ("yax", "xax")
        var = nc.createVariable(name, "d", dims)
        var[:] = interior(a)
        var.units = "m"
    nc.close()
    got = {n: np.full_like(a, -5.0) for n, a in f.items()}
    t = gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, got)
    assert t == dict(loop=77, julianday=2451545, secondsofday=3600, timestep=12.25)
    for n, a in f.items():
        assert np.array_equal(interior(got[n]), interior(a)), n
        halo = np.ones(a.shape[-2:], bool)
        halo[2:-2, 2:-2] = False
        assert np.all(got[n][..., halo] == -5.0), n  # the halos are kept


def test_round_trip_and_errors(tmp_path):
    f = fields(3)
    p = tmp_path / "rt.nc"
    gt.restart_write_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, f, loop=-3, timestep=-1.5)
    got = {n: np.zeros_like(a) for n, a in f.items()}
    t = gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, got)
    assert t["loop"] == -3 and t["timestep"] == -1.5
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


@pytest.mark.parametrize("missing", ["loop", "julianday", "secondsofday", "timestep"])
def test_a_file_without_a_required_scalar_is_rejected(tmp_path, missing):
    """open_restart_ncdf.F90:48-62 stops when one of the scalars is absent;
    so does the reader (also for a scalar of the wrong type)"""
    f = fields(5)
    for wrong_type in (False, True):
        p = tmp_path / f"m_{missing}_{wrong_type}.nc"
        nc = netcdf_file(str(p), "w")
        nc.createDimension("xax", IMAX - IMIN + 1)
        nc.createDimension("yax", JMAX - JMIN + 1)
        nc.createDimension("zax", KMAX + 1)
        for name, typ in (("loop", "i"), ("julianday", "i"), ("secondsofday", "i"),
                          ("timestep", "d")):
            if name == missing:
                if wrong_type:
                    nc.createVariable(name, "d" if typ == "i" else "i", ()).data[()] = 1
                continue
            nc.createVariable(name, typ, ()).data[()] = 1
        nc.createVariable("z", "d", ("yax", "xax"))[:] = interior(f["z"])
        nc.close()
        with pytest.raises(gt.GtError, match=missing):
            gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, {"z": f["z"].copy()})


def test_halo_file_reads_only_the_interior(tmp_path):
    """the __READ_HOT_HALOS_ build (read_restart_ncdf.F90:83-90) reads the
    interior of a halo-inclusive file and leaves the halos to the caller's
    halo update; it also takes an interior-sized file.  Without halos a
    halo-inclusive file does not match."""
    f = fields(9)
    ph = tmp_path / "halos.nc"
    pi = tmp_path / "interior.nc"
    gt.restart_write_host(ph, IMIN, IMAX, JMIN, JMAX, KMAX, f, with_halos=True)
    gt.restart_write_host(pi, IMIN, IMAX, JMIN, JMAX, KMAX, f, with_halos=False)
    for p in (ph, pi):
        got = {n: np.full_like(a, -7.0) for n, a in f.items()}
        gt.restart_read_host(p, IMIN, IMAX, JMIN, JMAX, KMAX, got, with_halos=True)
        for n, a in f.items():
            assert np.array_equal(interior(got[n]), interior(a)), (p.name, n)
            halo = np.ones(a.shape[-2:], bool)
            halo[2:-2, 2:-2] = False
            assert np.all(got[n][..., halo] == -7.0), (p.name, n)
    with pytest.raises(gt.GtError, match="xax"):
        gt.restart_read_host(ph, IMIN, IMAX, JMIN, JMAX, KMAX, {"z": f["z"].copy()})
