"""ctypes binding of the oracle (oracle/getm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Arrays are numpy float64 (kmax+1, NY, NX) or (NY, NX), C-contiguous, which
is the reference's Fortran memory order.  Functions modify their in/out
arguments in place, like the Fortran routines they restate.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libgetm_oracle.so"

NOADV, UPSTREAM, UPSTREAM_2DH, P2, SUPERBEE, MUSCL, P2_PDM, J7, FCT, P2_2DH = range(10)
NOSPLIT, FULLSPLIT, HALFSPLIT, HVSPLIT = range(4)
H_TAG, D_TAG, U_TAG, V_TAG, W_TAG = 10, 20, 31, 32, 33


class Dom(C.Structure):
    _fields_ = [("imin", C.c_int), ("imax", C.c_int), ("jmin", C.c_int), ("jmax", C.c_int),
                ("kmax", C.c_int), ("dx", C.c_double), ("dy", C.c_double),
                ("ard1", C.c_double)]


class AdvGrid(C.Structure):
    _fields_ = [("mask_uflux", C.c_void_p), ("mask_vflux", C.c_void_p),
                ("mask_uupdate", C.c_void_p), ("mask_vupdate", C.c_void_p),
                ("mask_finalise", C.c_void_p), ("az", C.c_void_p)]


class AdvGrids(C.Structure):
    _fields_ = [("H", AdvGrid), ("U", AdvGrid), ("V", AdvGrid), ("store", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} missing: run `python code_amd/build.py`")
        L = C.CDLL(str(LIB))
        vp, d, i = C.c_void_p, C.c_double, C.c_int
        D = C.POINTER(Dom)
        G = C.POINTER(AdvGrids)
        sig = {
            "or_adv_interfacial_reconstruction": (d, [i, d, d, d, d]),
            "or_masks_from_az": (i, [D, vp, vp, vp, vp]),
            "or_init_advection": (i, [D, vp, vp, vp, vp, G]),
            "or_free_advection": (None, [G]),
            "or_adv_split_u": (None, [D, d, vp, vp, vp, vp, vp, vp, d, i, d, vp, vp]),
            "or_adv_split_v": (None, [D, d, vp, vp, vp, vp, vp, vp, d, i, d, vp, vp]),
            "or_adv_split_w": (C.c_long, [D, d, vp, vp, vp, vp, vp, d, i, i, vp, i]),
            "or_adv_upstream_2dh": (None, [D, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp]),
            "or_do_advection": (i, [D, G, d, vp, vp, vp, vp, vp, vp, vp, i, i, d, i, vp, vp]),
            "or_do_advection_3d": (i, [D, G, d, vp, vp, vp, vp, vp, vp, vp, vp, i, i, i, d, i,
                                       i, vp, vp, C.POINTER(C.c_long)]),
            "or_eos_linear": (None, [D, vp, vp, vp, vp, d, d, d, d]),
            "or_rho_unesco80": (d, [d, d, i]),
            "or_rho_from_theta": (None, [d, d, d, C.POINTER(d), C.POINTER(d)]),
            "or_eosall_from_theta": (None, [d, d, d, C.POINTER(d), C.POINTER(d)]),
            "or_do_eqstate": (i, [D, vp, i, i, vp, vp, vp, vp, vp, vp, vp, d, d, d, d]),
            "or_ip_blumberg_mellor": (None, [D, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            "or_ip_shchepetkin_mcwilliams": (None, [D, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                                    vp, vp]),
            "or_substr_ini_press": (None, [D, C.POINTER(i), vp, vp, vp, vp]),
            "or_update_3d_halo": (None, [D, vp, vp, i]),
            "or_update_2d_halo": (None, [D, vp, vp, i, C.POINTER(C.c_int)]),
            "or_exchange_tiles": (None, [i, i, vp, vp, i, vp]),
            "or_decomp_1d": (None, [i, i, i, C.POINTER(i), C.POINTER(i)]),
            "or_budget": (None, [D, vp, vp, vp, C.POINTER(d), C.POINTER(d)]),
            "or_start_macro": (None, [D, i, d] + [vp] * 15),
            "or_coord_new": (vp, [D, i, d, d, d, i, d]),
            "or_coord_delete": (None, [vp]),
            "or_coordinates": (i, [D, vp, d, d, d] + [vp] * 16),
            "or_ww_momentum_3d": (None, [D, vp, d, vp, vp, vp, vp, vp]),
            "or_tridiagonal": (None, [i, i, i, vp, vp, vp, vp, vp]),
            "or_salinity_diffusion": (None, [D, vp, d, d, vp, vp, vp]),
            "or_temperature_diffusion": (None, [D, vp, d, d] + [vp] * 9 + [d, d, vp, vp]),
            "or_uv_advect_3d": (i, [D, G, d] + [vp] * 7 + [i, i, i, i, vp, vp]),
            "or_set_threads": (None, [i]),
            "or_get_threads": (i, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def P(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags.c_contiguous, "need C-contiguous numpy"
    return a.ctypes.data_as(C.c_void_p)


def reconstruct(scheme, cfl, fuu, fu, fd):
    return lib().or_adv_interfacial_reconstruction(scheme, cfl, fuu, fu, fd)


def rho_unesco80(T, S, nonnegsalt=False):
    return lib().or_rho_unesco80(T, S, int(nonnegsalt))


def rho_from_theta(s, th, p):
    """(dens0, densp) of eqstate.F90:338-423"""
    a, b = C.c_double(), C.c_double()
    lib().or_rho_from_theta(s, th, p, C.byref(a), C.byref(b))
    return a.value, b.value


def eosall_from_theta(s, th, p):
    """(rho_s, rho_th) of eqstate.F90:432-562"""
    a, b = C.c_double(), C.c_double()
    lib().or_eosall_from_theta(s, th, p, C.byref(a), C.byref(b))
    return a.value, b.value


def tridiagonal(kmax, fi, lt, au, bu, cu, du):
    """getm_tridiagonal (tridiagonal.F90:8-51) on arrays of kmax+1"""
    value = np.zeros(kmax + 1)
    arr = [np.ascontiguousarray(a, np.float64) for a in (au, bu, cu, du)]
    lib().or_tridiagonal(kmax, fi, lt, *(P(a) for a in arr), P(value))
    return value


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def decomp_1d(n, np_, m):
    s, e = C.c_int(), C.c_int()
    lib().or_decomp_1d(n, np_, m, C.byref(s), C.byref(e))
    return s.value, e.value


class Oracle:
    """The restatement on one domain (masks from az/au/av/ax as int32)."""

    def __init__(self, imax, jmax, kmax, dx, dy, az, au=None, av=None, ax=None, imin=1, jmin=1):
        self.dom = Dom(imin, imax, jmin, jmax, kmax, dx, dy, 1.0 / (dx * dy))
        self.NX, self.NY = imax - imin + 5, jmax - jmin + 5
        self.az = np.ascontiguousarray(az, np.int32)
        if au is None:
            au = np.zeros_like(self.az)
            av = np.zeros_like(self.az)
            ax = np.zeros_like(self.az)
            lib().or_masks_from_az(C.byref(self.dom), P(self.az), P(au), P(av), P(ax))
        self.au = np.ascontiguousarray(au, np.int32)
        self.av = np.ascontiguousarray(av, np.int32)
        self.ax = np.ascontiguousarray(ax, np.int32)
        self.g = AdvGrids()
        rc = lib().or_init_advection(C.byref(self.dom), P(self.az), P(self.au), P(self.av),
                                     P(self.ax), C.byref(self.g))
        assert rc == 0

    @classmethod
    def from_domain(cls, dom):
        return cls(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, dom.az, dom.au, dom.av,
                   dom.ax, imin=dom.imin, jmin=dom.jmin)

    def __del__(self):
        try:
            lib().or_free_advection(C.byref(self.g))
            if getattr(self, "_coord", None):
                lib().or_coord_delete(self._coord)
        except Exception:
            pass

    def grid(self, tag):
        return {H_TAG: self.g.H, D_TAG: self.g.H, W_TAG: self.g.H, U_TAG: self.g.U,
                V_TAG: self.g.V}[tag]

    def adv_split_u(self, dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH, tag=H_TAG):
        g = self.grid(tag)
        lib().or_adv_split_u(C.byref(self.dom), dt, P(f), P(fi), P(Di), P(adv), P(U), P(DU),
                             splitfac, scheme, AH, g.mask_uflux, g.mask_uupdate)

    def adv_split_v(self, dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH, tag=H_TAG):
        g = self.grid(tag)
        lib().or_adv_split_v(C.byref(self.dom), dt, P(f), P(fi), P(Di), P(adv), P(V), P(DV),
                             splitfac, scheme, AH, g.mask_vflux, g.mask_vupdate)

    def adv_split_w(self, dt, f, fi, hi, adv, ww, splitfac, scheme, tag, itersmax):
        g = self.grid(tag)
        return lib().or_adv_split_w(C.byref(self.dom), dt, P(f), P(fi), P(hi), P(adv), P(ww),
                                    splitfac, scheme, tag, g.az, itersmax)

    def adv_upstream_2dh(self, dt, f, fi, Di, adv, U, V, Dn, DU, DV, AH, tag=H_TAG):
        g = self.grid(tag)
        lib().or_adv_upstream_2dh(C.byref(self.dom), dt, P(f), P(fi), P(Di), P(adv), P(U),
                                  P(V), P(Dn), P(DU), P(DV), AH, g.az)

    def do_advection(self, dt, f, U, V, DU, DV, Do, Dn, split, scheme, AH, tag, Dires=None,
                     advres=None):
        return lib().or_do_advection(C.byref(self.dom), C.byref(self.g), dt, P(f), P(U), P(V),
                                     P(DU), P(DV), P(Do), P(Dn), split, scheme, AH, tag,
                                     P(Dires), P(advres))

    def do_advection_3d(self, dt, f, uu, vv, ww, hu, hv, ho, hn, split, hscheme, vscheme, AH,
                        tag, adv_ver_iterations=1, hires=None, advres=None):
        nw = C.c_long(0)
        rc = lib().or_do_advection_3d(C.byref(self.dom), C.byref(self.g), dt, P(f), P(uu),
                                      P(vv), P(ww), P(hu), P(hv), P(ho), P(hn), split, hscheme,
                                      vscheme, AH, tag, adv_ver_iterations, P(hires),
                                      P(advres), C.byref(nw))
        return rc, nw.value

    def eos_linear(self, T, S, rho, buoy, T0, S0, dtr0, dsr0):
        lib().or_eos_linear(C.byref(self.dom), P(T), P(S), P(rho), P(buoy), T0, S0, dtr0, dsr0)

    def do_eqstate(self, method, T, S, hn, rho, buoy, alpha=None, beta=None, T0=10.0,
                   S0=33.75, dtr0=-0.17, dsr0=0.78, nonnegsalt=False):
        """do_eqstate (eqstate.F90:94-271); the defaults are the reference's
        namelist defaults (FP32 literals, rounded here)."""
        f32 = lambda v: float(np.float32(v))
        return lib().or_do_eqstate(C.byref(self.dom), P(self.az), method, int(nonnegsalt),
                                   P(T), P(S), P(hn), P(rho), P(buoy), P(alpha), P(beta),
                                   f32(T0), f32(S0), f32(dtr0), f32(dsr0))

    # ---- input producers (m3d.F90:446-548) ----
    def start_macro(self, M, min_depth, H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn,
                    Dn, Dun, Dvn, Uint, Vint):
        lib().or_start_macro(C.byref(self.dom), M, min_depth,
                             *(P(a) for a in (H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn,
                                              Dn, Dun, Dvn, Uint, Vint)))

    def coordinates(self, dt, cord_relax, maxdepth, H, HU, HV, sseo, ssen, ssuo, ssun, ssvo,
                    ssvn, ho, hn, huo, hun, hvo, hvn, zc=None):
        """coordinates(); the first call initialises (set_coordinates
        chooses the kind and its namelist values)"""
        return lib().or_coordinates(C.byref(self.dom), self._coord, dt, cord_relax, maxdepth,
                                    *(P(a) for a in (H, HU, HV, sseo, ssen, ssuo, ssun, ssvo,
                                                     ssvn, ho, hn, huo, hun, hvo, hvn, zc)))

    def set_coordinates(self, vert_cord, ddu=0.0, ddl=0.0, d_gamma=20.0, gamma_surf=True,
                        min_depth=0.5):
        if getattr(self, "_coord", None):
            lib().or_coord_delete(self._coord)
        self._coord = lib().or_coord_new(C.byref(self.dom), vert_cord, ddu, ddl, d_gamma,
                                         int(gamma_surf), min_depth)

    def ww_momentum_3d(self, dt, uu, vv, ho, hn, ww):
        lib().or_ww_momentum_3d(C.byref(self.dom), P(self.az), dt, P(uu), P(vv), P(ho), P(hn),
                                P(ww))

    def uv_advect_3d(self, dt, uu, vv, ww, ho, hn, hun, hvn, split, hscheme, vscheme, uuEx,
                     vvEx, adv_ver_iterations=1):
        return lib().or_uv_advect_3d(C.byref(self.dom), C.byref(self.g), dt,
                                     *(P(a) for a in (uu, vv, ww, ho, hn, hun, hvn)), split,
                                     hscheme, vscheme, adv_ver_iterations, P(uuEx), P(vvEx))

    # ---- vertical diffusion (temperature.F90:463-551, salinity.F90:346-393) ----
    def salinity_diffusion(self, dt, cnpar, hn, nuh, S):
        lib().or_salinity_diffusion(C.byref(self.dom), P(self.az), dt, cnpar, P(hn), P(nuh),
                                    P(S))

    def temperature_diffusion(self, dt, cnpar, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T,
                              swr_bot_refl_frac=-1.0, swr_min_bot_frac=0.01):
        f32 = lambda v: float(np.float32(v))
        lib().or_temperature_diffusion(C.byref(self.dom), P(self.az), dt, cnpar, P(hn), P(nuh),
                                       P(swr), P(shf), P(A), P(g1), P(g2), P(dry_z), P(S),
                                       swr_bot_refl_frac, f32(swr_min_bot_frac), P(rad), P(T))

    def ip_blumberg_mellor(self, H, hn, hun, hvn, buoy, zz, idpdx, idpdy):
        lib().or_ip_blumberg_mellor(C.byref(self.dom), P(self.az), P(self.au), P(self.av),
                                    P(H), P(hn), P(hun), P(hvn), P(buoy), P(zz), P(idpdx),
                                    P(idpdy))

    def ip_shchepetkin_mcwilliams(self, H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy):
        lib().or_ip_shchepetkin_mcwilliams(C.byref(self.dom), P(self.az), P(self.au),
                                           P(self.av), P(H), P(hn), P(hun), P(hvn), P(buoy),
                                           P(sseo), P(zz), P(idpdx), P(idpdy))

    def do_internal_pressure(self, ip_method, H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy,
                             substr=None):
        """do_internal_pressure (internal_pressure.F90:181-274) for methods 1
        and 6; substr: None, or a dict holding the SUBSTR_INI_PRESS state
        (keys first, idpdx0, idpdy0; created on the first call)"""
        if ip_method == 1:
            self.ip_blumberg_mellor(H, hn, hun, hvn, buoy, zz, idpdx, idpdy)
        elif ip_method == 6:
            self.ip_shchepetkin_mcwilliams(H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy)
        else:
            raise ValueError(f"ip_method {ip_method} is not restated")
        if substr is not None:
            if "first" not in substr:
                substr.update(first=1, idpdx0=np.zeros_like(idpdx), idpdy0=np.zeros_like(idpdy))
            fl = C.c_int(substr["first"])
            lib().or_substr_ini_press(C.byref(self.dom), C.byref(fl), P(idpdx), P(idpdy),
                                      P(substr["idpdx0"]), P(substr["idpdy0"]))
            substr["first"] = fl.value

    def update_3d_halo(self, f1, f2, mirror=-1):
        lib().or_update_3d_halo(C.byref(self.dom), P(f1), P(f2), mirror)

    def update_2d_halo(self, f1, f2, mirror, state):
        st = C.c_int(state)
        lib().or_update_2d_halo(C.byref(self.dom), P(f1), P(f2), mirror, C.byref(st))
        return st.value

    def budget(self, h, f=None):
        m, v = C.c_double(), C.c_double()
        lib().or_budget(C.byref(self.dom), P(self.az), P(h), P(f), C.byref(m), C.byref(v))
        return m.value, v.value


def exchange_tiles(px, py, ni, nj, nk, tiles):
    n = px * py
    ni_a = (C.c_int * n)(*ni)
    nj_a = (C.c_int * n)(*nj)
    tp = (C.c_void_p * n)(*[t.ctypes.data for t in tiles])
    lib().or_exchange_tiles(px, py, ni_a, nj_a, nk, tp)
