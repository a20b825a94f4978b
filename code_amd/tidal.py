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

    def snapshot(self):
        """the state a macro step starts from: the time, the elevation, the
        2-D fields of start_macro and the layer heights the coordinates relax
        from (hn, hun, hvn; ho, huo, hvo and zc are rewritten from them)"""
        return (self.t, self.z.clone(), {n: v.clone() for n, v in self.e.items()},
                {n: self.h[n].clone() for n in ("hn", "hun", "hvn")})

    def restore(self, snap):
        """back to a snapshot: the macro steps that follow repeat bit for bit"""
        t, z, e, h = snap
        self.t = t
        self.z.copy_(z)
        for n, v in e.items():
            self.e[n].copy_(v)
        for n, v in h.items():
            self.h[n].copy_(v)

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


STEP_INPUTS = ("uu", "vv", "ww", "hu", "hv", "ho")


class TidalAdjointWindow:
    """The FP32 reverse mode of the tracers over nsteps macro steps of the
    tide, with memory bounded for the full 2000x1500x50 grid on one GPU.

    The per-step inputs (six FP64 fields, 7.4 GB a step at that size) are
    never all kept: the forward sweep keeps, every `seg` steps, the tide's
    state (snapshot: 2-D fields and three layer-height fields) and the tracer
    set; the reverse sweep restores each snapshot, regenerates the inputs of
    that segment's steps through the producers (bit for bit, the producers
    are deterministic) and hands them with the segment's tracer checkpoint to
    gt_adv3d_adjoint_seq_f32, `batch` tracers at a time (the tracers are
    independent given the inputs, so batching changes no bit), the library
    itself keeping a checkpoint every `inner_ckpt` steps inside the segment.
    lam_final -> lam0 equals one gt_adv3d_adjoint_seq_f32 call over all
    nsteps steps with every step's inputs kept."""

    def __init__(self, tide, ctx, nsteps=24, seg=6, inner_ckpt=2, batch=6,
                 scheme=None, stream=None):
        from . import transport as gt
        self.tide, self.ctx = tide, ctx
        self.nsteps, self.seg, self.inner_ckpt, self.batch = nsteps, seg, inner_ckpt, batch
        self.scheme = gt.P2_PDM if scheme is None else scheme
        self.stream = stream
        self.ckpts = []

    def forward(self, f0, f_out, hires=None):
        """nsteps batched FP32 steps from f0 (left untouched); the result in
        f_out; keeps the checkpoints (f0 itself is the first tracer set)"""
        from . import transport as gt
        nt = f0.shape[0]
        self.ckpts = []
        cur, nxt = f0, f_out
        spare = None
        for t in range(self.nsteps):
            if t % self.seg == 0:
                self.ckpts.append((self.tide.snapshot(), f0 if t == 0 else cur.clone()))
            dd, _ = self.tide.macro_step()
            gt.adv3d_batch([dict(dd, ctx=self.ctx, f_in=cur, f_out=nxt, hires=hires)], nt,
                           self.tide.dt, self.scheme, self.scheme, stream=self.stream)
            if t == 0:
                # f0 stays the caller's: the next steps ping-pong between
                # f_out and a work set
                spare = f0.new_empty(f0.shape)
                cur, nxt = nxt, spare
            else:
                cur, nxt = nxt, cur
        if cur is not f_out:
            f_out.copy_(cur)
        del spare
        return f_out

    def reverse(self, lam):
        """lam (the adjoint of the final tracer set) is replaced by the
        adjoint of f0, segment by segment from the last"""
        nt = lam.shape[0]
        for s in range(len(self.ckpts) - 1, -1, -1):
            snap, fck = self.ckpts[s]
            self.tide.restore(snap)
            t0 = s * self.seg
            t1 = min(self.nsteps, t0 + self.seg)
            seq = []
            for _ in range(t0, t1):
                dd, _ = self.tide.macro_step()
                seq.append(tuple(dd[n].clone() for n in STEP_INPUTS))
            for b0 in range(0, nt, self.batch):
                b1 = min(nt, b0 + self.batch)
                self.ctx.adv3d_adjoint_seq(b1 - b0, self.inner_ckpt, self.tide.dt, fck[b0:b1],
                                           lam[b0:b1], lam[b0:b1], seq, self.scheme,
                                           self.scheme)
            del seq
        return lam

    def release(self):
        self.ckpts = []
