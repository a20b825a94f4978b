"""Verification harness for the reverse mode: torch.autograd.Function
wrappers around the HIP forward (gt_do_advection_3d) and its hand-written
adjoint (gt_adv3d_adjoint, gt_ip_blumberg_mellor_adjoint,
gt_eos_linear_adjoint).  Used by the tests only; the library itself does not
depend on autograd."""
import torch

from code_amd import inputs
from code_amd import transport as gt


class AdvectSteps(torch.autograd.Function):
    """f_final = (HALFSPLIT step)^n (f0), velocities and heights fixed"""

    @staticmethod
    def forward(ctx, f0, gctx, fields, nsteps, ckpt, hs, vs, dt):
        f = f0.detach().clone()
        uu, vv, ww, hu, hv, ho, hn = fields
        for _ in range(nsteps):
            gctx.do_advection_3d(dt, f, uu, vv, ww, hu, hv, ho, hn, gt.HALFSPLIT, hs, vs, -1.0,
                                 gt.H_TAG)
        gctx.synchronize()
        ctx.save_for_backward(f0.detach().clone())
        ctx.args = (gctx, fields, nsteps, ckpt, hs, vs, dt)
        return f

    @staticmethod
    def backward(ctx, lam):
        (f0,) = ctx.saved_tensors
        gctx, fields, nsteps, ckpt, hs, vs, dt = ctx.args
        uu, vv, ww, hu, hv, ho, hn = fields
        lam0 = torch.empty_like(f0)
        gctx.adv3d_adjoint(nsteps, ckpt, dt, f0, lam.contiguous(), lam0, uu, vv, ww, hu, hv, ho,
                           hs, vs)
        gctx.synchronize()
        return lam0, None, None, None, None, None, None, None


class PressureFromT(torch.autograd.Function):
    """idpdx, idpdy of Blumberg-Mellor from T (linear EOS, S fixed)"""

    @staticmethod
    def forward(ctx, T, S, gctx, H, hn, hun, hvn):
        consts = (inputs.EOS_T0, inputs.EOS_S0, inputs.EOS_DTR0, inputs.EOS_DSR0)
        buoy = torch.zeros_like(T)
        zz, ix, iy = (torch.zeros_like(T) for _ in range(3))
        gctx.eos_linear(T.detach(), S, None, buoy, *consts)
        gctx.ip_blumberg_mellor(H, hn, hun, hvn, buoy, zz, ix, iy)
        gctx.synchronize()
        ctx.args = (gctx, hun, hvn, zz)
        return ix, iy

    @staticmethod
    def backward(ctx, gx, gy):
        gctx, hun, hvn, zz = ctx.args
        bbar = torch.zeros_like(gx)
        gctx.ip_blumberg_mellor_adjoint(hun, hvn, zz, gx.contiguous(), gy.contiguous(), bbar)
        Tbar = torch.zeros_like(gx)
        gctx.eos_linear_adjoint(bbar, Tbar, None, inputs.EOS_DTR0, inputs.EOS_DSR0)
        gctx.synchronize()
        return Tbar, None, None, None, None, None, None
