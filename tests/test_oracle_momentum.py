"""Pins of the oracle's uv_advect_3d (momentum advection, a caller of the
transport path), CPU.

Closed forms on an all-water domain with uniform layers and no vertical
motion (interior points, away from the halo where the stencils are cut):
* u varying linearly across the flow, u = a*j, carried by a uniform
  transport V0*h along j: the advection term of u is V0*h*a/dy for every
  scheme (the flux difference of a linear field under a uniform transport;
  the limiters all give phi(1) = 1), and the V part, whose velocity is
  uniform, gives 0;
* the mirror case for v.
"""
import numpy as np
import pytest

from code_amd import inputs
from oracle import pyoracle as po
from tests.test_gpu_parity import random_domain

SCHEMES = [po.UPSTREAM, po.P2, po.SUPERBEE, po.MUSCL, po.P2_PDM]


def uniform_case(kmax=4):
    dom = random_domain(24, 22, kmax, 3, "open")
    orc = po.Oracle.from_domain(dom)
    sh = dom.shape3
    h = 2.0
    hn = np.full(sh, h)
    return dom, orc, sh, h, hn


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("split", [po.HALFSPLIT, po.FULLSPLIT])
def test_u_sheared_by_uniform_cross_transport(scheme, split):
    dom, orc, sh, h, hn = uniform_case()
    a, V0 = 0.01, 0.3
    jj = np.arange(dom.NY)[None, :, None].astype(float)
    uu = np.broadcast_to(h * a * jj, sh).copy()
    vv = np.full(sh, V0 * h)
    ww = np.zeros(sh)
    uuEx = np.full(sh, 7.0)
    vvEx = np.full(sh, 7.0)
    rc = orc.uv_advect_3d(10.0, uu, vv, ww, hn, hn, hn, hn, split, scheme, scheme, uuEx, vvEx)
    assert rc == 0
    inner = (slice(1, None), slice(5, -5), slice(5, -5))
    assert np.allclose(uuEx[inner], V0 * h * a / dom.dy, rtol=1e-9, atol=0)
    assert np.allclose(vvEx[inner], 0.0, rtol=0, atol=1e-15)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_v_sheared_by_uniform_cross_transport(scheme):
    dom, orc, sh, h, hn = uniform_case()
    b, U0 = -0.02, 0.25
    ii = np.arange(dom.NX)[None, None, :].astype(float)
    vv = np.broadcast_to(h * b * ii, sh).copy()
    uu = np.full(sh, U0 * h)
    ww = np.zeros(sh)
    uuEx, vvEx = np.zeros(sh), np.zeros(sh)
    rc = orc.uv_advect_3d(10.0, uu, vv, ww, hn, hn, hn, hn, po.HALFSPLIT, scheme, scheme, uuEx,
                          vvEx)
    assert rc == 0
    inner = (slice(1, None), slice(5, -5), slice(5, -5))
    assert np.allclose(vvEx[inner], U0 * h * b / dom.dx, rtol=1e-9, atol=0)
    assert np.allclose(uuEx[inner], 0.0, rtol=0, atol=1e-15)


def test_uv_advect_errors_and_noadv():
    dom, orc, sh, h, hn = uniform_case()
    z = np.zeros(sh)
    uuEx = np.full(sh, 3.0)
    assert orc.uv_advect_3d(10.0, z, z, z, hn, hn, hn, hn, po.HALFSPLIT, po.J7, po.J7, uuEx,
                            uuEx.copy()) == 3  # OR_EINVAL_SCHEME
    rc = orc.uv_advect_3d(10.0, z, z, z, hn, hn, hn, hn, po.HALFSPLIT, po.NOADV, po.NOADV, uuEx,
                          uuEx.copy())
    assert rc == 0 and np.all(uuEx == 3.0)  # nothing to do: untouched
