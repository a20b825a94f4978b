"""CPU checks of the boundary: the C-ABI library loads and exports every
entry point include/getm_transport.h declares (no compute calls: there is no
GPU here), the Fortran binding compiles, and the host-side mask logic matches
the oracle's."""
import re
from pathlib import Path

import numpy as np

from code_amd import build, inputs
from code_amd import transport as gt
from oracle import pyoracle as po

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "getm_transport.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = gt.load_library()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(gt.EXPORTS) == names


def test_version_and_error_string():
    lib = gt.load_library()
    assert lib.gt_version() >= 1
    assert isinstance(lib.gt_last_error(), bytes)


def test_library_is_gfx950_code_object():
    data = (build.BUILD / build.LIB_NAME).read_bytes()
    assert b"gfx950" in data


def test_fortran_binding_builds():
    obj = build.build_fortran()
    assert obj is not None and obj.exists()
    assert (build.BUILD / "fortran" / "getm_transport_gpu.mod").exists()


def test_masks_match_domain_rules():
    dom = inputs.make_domain(17, 13, 3, 4, walls=2)
    # add an open-boundary cell (az=2) to exercise the 2/3 codes
    az = dom.az.copy()
    az[5, 3] = 2
    az[5, 4] = 2
    au, av, ax = inputs.masks_from_az(az)
    orc = po.Oracle(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, az)
    assert np.array_equal(au, orc.au)
    assert np.array_equal(av, orc.av)
    assert np.array_equal(ax, orc.ax)
    assert (au == 2).any() and (au == 3).any()
