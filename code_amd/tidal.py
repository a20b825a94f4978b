"""Time-varying forcing of the 2000x1500x50 tidal case on one GPU's share.

GETM's 2-D mode (outside the transport path) is stood in for by a prescribed
M2 seiche of the depth-integrated transport,
U = umax * Href * sin(pi s) * cos(omega t), V = 0, with s the position across
the water span (zero on the walls) and Href the smallest water depth at the
U points: the velocity U/HU peaks at umax where the water is shallowest, and
the divergence, hence the elevation change and the vertical velocity, comes
from the seiche's shape alone (a transport proportional to the local depth
would pump the bathymetry's roughness into the vertical), advanced over the
M micro steps of each
macro step with the elevation update of sealevel.F90:92-94 and the transport
sum of m2d.F90:471.  Everything else runs through the library on the
device, in integrate_3d's order (m3d.F90:446-548): start_macro, coordinates
(general), the layer transports uu_k = Uint/M * hun_k/Dun, ww_momentum_3d.
The elevation change equals the divergence of the mean transport, so the
layer-height trajectory of the advection closes on hn to round-off.
"""
from __future__ import annotations

import math

import numpy as np

M2_PERIOD = 44714.0


class TidalShare:
    """Forcing and input producers of one subdomain (device tensors)."""

    def __init__(self, dom, ctx, device, M=30, dt=300.0, umax=1.0, d_gamma=20.0,
                 min_depth=0.5, phase=0.0):
        import torch
        self.dom, self.ctx, self.dev = dom, ctx, device
        self.M, self.dt, self.dt2d = M, dt, dt / M
        self.min_depth = min_depth
        self.omega = 2.0 * math.pi / M2_PERIOD
        self.t = phase
        D = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(device)  # noqa
        self.H, self.HU, self.HV = D(dom.H), D(dom.HU), D(dom.HV)
        wet_cols = np.where(dom.az.max(axis=0) == 1)[0]
        i0, i1 = wet_cols.min(), wet_cols.max()
        ii = np.arange(dom.NX, dtype=np.float64)
        s = (ii + 0.5 - (i0 - 0.5)) / float(i1 - i0 + 1)  # face i+1/2 across the water span
        shape = np.clip(np.sin(math.pi * s), 0.0, None)[None, :] * (dom.au >= 1)
        self.Href = float(dom.HU[dom.au >= 1].min())
        self.Ushape = D(umax * self.Href * shape)
        self.wet = D((dom.az == 1).astype(np.float64))
        sh2 = dom.az.shape
        z2 = lambda: torch.zeros(sh2, dtype=torch.float64, device=device)  # noqa: E731
        self.z = z2()
        self.e = {n: z2() for n in ("sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn", "Dn", "Dun",
                                    "Dvn", "Uint", "Vint")}
        sh3 = dom.shape3
        z3 = lambda: torch.zeros(sh3, dtype=torch.float64, device=device)  # noqa: E731
        self.h = {n: z3() for n in ("ho", "hn", "huo", "hun", "hvo", "hvn", "zc")}
        self.uu, self.vv, self.ww = z3(), z3(), z3()
        ctx.set_coordinates(3, ddu=1.0, ddl=1.0, d_gamma=d_gamma, gamma_surf=True,
                            min_depth=min_depth)

    def micro_steps(self):
        """M micro steps of the stand-in 2-D mode: U at the micro time, the
        elevation update of sealevel.F90:92-94, Uint += U (m2d.F90:471)"""
        dom = self.dom
        self.e["Uint"].zero_()
        self.e["Vint"].zero_()
        for _ in range(self.M):
            U = self.Ushape * math.cos(self.omega * self.t)
            div = (U * dom.dy - U.roll(1, 1) * dom.dy) * dom.ard1
            self.z -= self.dt2d * div * self.wet
            self.e["Uint"] += U
            self.t += self.dt2d

    def macro_step(self):
        """the micro steps and the producers of one macro step; returns the
        device inputs of the advection (uu, vv, ww, hun, hvn, ho) and hn"""
        self.micro_steps()
        return self.producers()

    def producers(self):
        """start_macro, coordinates, layer transports, ww_momentum_3d"""
        c, e, h = self.ctx, self.e, self.h
        c.start_macro(self.M, self.min_depth, self.H, self.HU, self.HV, self.z,
                      *(e[n] for n in ("sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn", "Dn",
                                       "Dun", "Dvn", "Uint", "Vint")))
        c.coordinates(self.dt, 0.0, 50.0, self.H, self.HU, self.HV,
                      *(e[n] for n in ("sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn")),
                      *(h[n] for n in ("ho", "hn", "huo", "hun", "hvo", "hvn", "zc")))
        # the layer transports: the mean transport split by hun_k / Dun
        Dun = e["Dun"]
        frac = torch_where_pos(Dun)
        self.uu.copy_(h["hun"] * (e["Uint"] * frac)[None])
        self.uu[0].zero_()
        self.vv.zero_()
        self.ww.zero_()
        c.ww_momentum_3d(self.dt, self.uu, self.vv, h["ho"], h["hn"], self.ww)
        return dict(uu=self.uu, vv=self.vv, ww=self.ww, hu=h["hun"], hv=h["hvn"], ho=h["ho"]), \
            h["hn"]


def torch_where_pos(D):
    """1/D where D > 0, else 0"""
    import torch
    return torch.where(D > 0, 1.0 / torch.where(D > 0, D, torch.ones_like(D)),
                       torch.zeros_like(D))


def initial_tracers(dom, device, nt=18, seed=5):
    """T (warmer above), S (a horizontal gradient), nt-3 seeded Gaussian
    blobs and one constant tracer, FP32, level 0 zero"""
    import torch
    kmax = dom.kmax
    f = torch.zeros((nt,) + dom.shape3, dtype=torch.float32, device=device)
    k = torch.arange(kmax + 1, device=device, dtype=torch.float64)[:, None, None]
    x = torch.arange(dom.NX, device=device, dtype=torch.float64)[None, None, :] / dom.NX
    y = torch.arange(dom.NY, device=device, dtype=torch.float64)[None, :, None] / dom.NY
    f[0] = (5.0 + 15.0 * k / kmax + 0.5 * torch.sin(6.0 * x)).float()
    f[1] = (25.0 + 5.0 * torch.sin(math.pi * x)).float().expand(kmax + 1, dom.NY, -1)
    rng = np.random.default_rng(seed)
    for q in range(2, nt - 1):
        cx, cy, sg, amp = rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), \
            rng.uniform(0.02, 0.07), rng.uniform(0.0, 1.0)
        r2 = ((x - cx) ** 2 + (y - cy) ** 2) / (2 * sg * sg)
        f[q] = (amp * torch.exp(-r2)).float().expand(kmax + 1, -1, -1)
    f[nt - 1] = 1.0
    f[:, 0] = 0.0
    return f
