"""The decomposition itself, on the CPU with the oracle: the parallel
build's HALFSPLIT sequence (tests/decomposed.py) over 2 x 2 tiles of uneven
size equals the undecomposed step bit for bit on a closed basin (land in
the global ghost ring, islands inside) for every scheme and tag.  With water
in the global ghost ring it does not, and only where the reference's TWOD
pattern leaves a tile's ghost cells stale: the HALO x HALO corner of a tile
where its ghost strip along an inner tile edge meets the global ghost ring
(no diagonal neighbour sends it, halo_mpi.F90:1108-1115), whose values reach
the interior cells next to that corner only."""
import numpy as np
import pytest

from code_amd import transport as gt
from oracle import pyoracle as po
from tests import decomposed as dc
from tests.test_gpu_parity import random_domain, random_fields


def _run(dom, F, scheme, tag):
    dt, AH = F["dt"], -1.0
    want = F["f"].copy()
    rc, _ = po.Oracle.from_domain(dom).do_advection_3d(
        dt, want, F["uu"], F["vv"], F["ww"], F["hu"], F["hv"], F["ho"], F["hn"], gt.HALFSPLIT,
        scheme, scheme, AH, tag)
    assert rc == 0
    tiles = [dict(orc=po.Oracle.from_domain(td), sl=sl, hi=fl["ho"].copy(),
                  adv=np.zeros_like(fl["ho"]), ni=td.imax, nj=td.jmax, **fl)
             for td, sl, fl in dc.tiles_of(dom, F)]

    def horizontal(u):
        for t in tiles:
            for k in range(1, dom.kmax + 1):
                f = t["f"][k].copy()
                fi = f.copy()
                vel, h = (t["uu"][k], t["hu"][k]) if u else (t["vv"][k], t["hv"][k])
                fn = t["orc"].adv_split_u if u else t["orc"].adv_split_v
                hi, adv = t["hi"][k].copy(), t["adv"][k].copy()
                fn(dt, f, fi, hi, adv, np.ascontiguousarray(vel), np.ascontiguousarray(h), 0.5,
                   scheme, AH, tag)
                t["f"][k], t["hi"][k], t["adv"][k] = fi, hi, adv

    def w_step():
        for t in tiles:
            fi = t["f"].copy()
            t["orc"].adv_split_w(dt, t["f"].copy(), fi, t["hi"], t["adv"], t["ww"], 1.0, scheme,
                                 tag, 1)
            t["f"] = fi

    def exchange():
        po.exchange_tiles(2, 2, [t["ni"] for t in tiles], [t["nj"] for t in tiles],
                          dom.kmax + 1, [t["f"] for t in tiles])

    dc.halfsplit(scheme, tag, lambda: horizontal(True), lambda: horizontal(False), w_step,
                 exchange)
    return want, tiles


@pytest.mark.parametrize("kind", ["walls", "closed_islands"])
@pytest.mark.parametrize("scheme,tag", dc.CASES)
def test_decomposed_equals_undecomposed_on_a_closed_basin(kind, scheme, tag):
    dom = random_domain(38, 33, 5, 21, "walls") if kind == "walls" else dc.closed_islands()
    F = random_fields(dom, 300, plateaus=True)
    want, tiles = _run(dom, F, scheme, tag)
    for q, t in enumerate(tiles):
        ref = want[(slice(None),) + t["sl"]]
        inner = (slice(None), slice(2, t["nj"] + 2), slice(2, t["ni"] + 2))
        assert np.array_equal(t["f"][inner], ref[inner]), q


def test_open_boundary_differs_only_next_to_the_unsent_corners():
    dom = random_domain(38, 33, 5, 21, "open")
    F = random_fields(dom, 300, plateaus=True)
    want, tiles = _run(dom, F, gt.P2_PDM, gt.H_TAG)
    NX, NY = dom.imax + 4, dom.jmax + 4
    ndiff = 0
    for t in tiles:
        ys, xs = t["sl"]
        ref = want[:, ys, xs]
        d = t["f"] != ref
        d[:, :2, :] = d[:, -2:, :] = False  # the tile's own ghost ring
        d[:, :, :2] = d[:, :, -2:] = False
        for k, j, i in np.argwhere(d):
            gj, gi = j + ys.start, i + xs.start  # global array indices
            near_global_x = gi <= 3 or gi >= NX - 4
            near_global_y = gj <= 3 or gj >= NY - 4
            near_tile_x = i <= 3 or i >= t["ni"] + 0
            near_tile_y = j <= 3 or j >= t["nj"] + 0
            assert (near_global_x and near_tile_y) or (near_global_y and near_tile_x), (k, gj, gi)
            ndiff += 1
    assert ndiff > 0
