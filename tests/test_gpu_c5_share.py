"""The tidal 2000x1500x50 case on one GPU's share (1000x375x50, general
coordinates, 18 FP32 tracers with FP64 budgets), macro steps with the input
producers on the device (code_amd/tidal.py): size-independent properties at
full size.

* the layer-height trajectory of the advection closes on the coordinates'
  hn (max |hires - hn| <= 1e-12 max hn): start_macro, coordinates and
  ww_momentum_3d produce volume-consistent inputs from the tidal transports;
* every tracer's FP64 mass budget sum(h f A) is conserved to the FP32
  rounding of the stored tracer (<= 1e-6 relative) while the volumes change;
* the horizontal and vertical Courant numbers stay below 1 (the batched
  step does not sub-cycle in the vertical) and the surface moves by less
  than 2 m;
* a constant tracer stays constant (within FP32 rounding); T makes no new
  extrema (P2-PDM is TVD at CFL < 1);
* FP32 T against the FP64 path on the same inputs: max error <= 1e-5 of max|T|.
"""
import pytest
import torch

from code_amd import inputs, tidal
from code_amd import transport as gt

pytestmark = pytest.mark.gpu


def test_c5_share_tidal_macro_steps(device):
    dom = inputs.make_domain(1000, 375, 50, 5, hmin=10.0, hmax=50.0)
    ctx = gt.Context.from_domain(dom)
    tide = tidal.TidalShare(dom, ctx, device, M=30, dt=300.0, umax=1.0)
    nt = 18
    f = tidal.initial_tracers(dom, device, nt)
    g = torch.empty_like(f)
    T64 = f[0:1].double().clone()
    T64b = torch.empty_like(T64)
    wet = torch.from_numpy(dom.az == 1).to(device)
    wet3 = wet[None].expand(dom.kmax, -1, -1)
    Tmin, Tmax = f[0, 1:][wet3].min().item(), f[0, 1:][wet3].max().item()
    hires = torch.zeros(dom.shape3, dtype=torch.float64, device=device)
    hires64 = torch.zeros_like(hires)
    zmax = 0.0
    for step in range(4):
        dev_in, hn = tide.macro_step()
        zmax = max(zmax, tide.z[wet].abs().max().item())
        h = dev_in["ho"]
        cw = (dev_in["ww"][1:-1].abs() * tide.dt / (0.5 * (h[1:-1] + h[2:])))[:, wet]
        hu = dev_in["hu"][1:]
        cu = torch.where(hu > 0, dev_in["uu"][1:].abs() / torch.where(hu > 0, hu, 1.0), 0.0)
        assert cw.max().item() < 1.0 and cu.max().item() * tide.dt / dom.dx < 1.0, step
        m0 = [ctx.budget(dev_in["ho"], f[q])[0] for q in range(nt)]
        gt.adv3d_batch([dict(ctx=ctx, f_in=f, f_out=g, hires=hires, **dev_in)], nt, tide.dt,
                       gt.P2_PDM, gt.P2_PDM)
        gt.adv3d_batch([dict(ctx=ctx, f_in=T64, f_out=T64b, hires=hires64, **dev_in)], 1,
                       tide.dt, gt.P2_PDM, gt.P2_PDM)
        for hr in (hires, hires64):
            err_h = ((hr - hn)[1:][wet3]).abs().max().item() / hn[1:][wet3].max().item()
            assert err_h <= 1e-12, (step, err_h)
        for q in range(nt):
            m1 = ctx.budget(hn, g[q])[0]
            assert abs(m1 - m0[q]) <= 1e-6 * abs(m0[q]), (step, q, m0[q], m1)
        c = g[nt - 1, 1:][wet3]
        assert (c - 1.0).abs().max().item() <= 1e-6, step
        Tg = g[0, 1:][wet3]
        assert Tg.min().item() >= Tmin - 1e-4 and Tg.max().item() <= Tmax + 1e-4, step
        f, g = g, f
        T64, T64b = T64b, T64
    assert 0.05 < zmax < 2.0  # the tide moved the surface
    e = (f[0, 1:].double() - T64[0, 1:])[wet3].abs().max().item() / T64[0, 1:][wet3].abs().max().item()
    assert e <= 1e-5, e
