"""Python binding of the C-ABI in include/getm_transport.h.

This is the host-side mirror used by tests and the benchmark: a thin ctypes
layer over ``libgetm_transport.so`` whose methods keep the argument lists of
the reference routines (getm-model/code) they replace.  Device memory and
streams come from PyTorch; every call goes to the HIP library, and loading
fails loudly if the library has not been built.

Fields are torch tensors of shape (kmax+1, NY, NX) (3-D) or (NY, NX) (one
slice), C-contiguous, so memory order equals the reference's Fortran order
f(imin-2:imax+2, jmin-2:jmax+2, 0:kmax).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import build as _build

# ids (advection.F90:55-73, advection_3d.F90:42, halo_zones.F90:24-28)
NOADV, UPSTREAM, UPSTREAM_2DH, P2, SUPERBEE, MUSCL, P2_PDM, J7, FCT, P2_2DH = range(10)
NOSPLIT, FULLSPLIT, HALFSPLIT, HVSPLIT = range(4)
H_TAG, D_TAG, U_TAG, V_TAG, W_TAG = 10, 20, 31, 32, 33
BLUMBERG_MELLOR, SHCHEPETKIN_MCWILLIAMS = 1, 6

GT_ERRORS = {1: "EINVAL_TAG", 2: "EINVAL_SPLIT", 3: "EINVAL_SCHEME",
             4: "ESCHEME_SPLIT", 5: "ENOMEM", 6: "EHIP", 7: "EARG",
             8: "ENOTIMPL"}

# every symbol include/getm_transport.h declares
EXPORTS = [
    "gt_last_error", "gt_version", "gt_create", "gt_destroy", "gt_set_stream",
    "gt_synchronize", "gt_set_adv_ver_iterations", "gt_warnings",
    "gt_adv_interfacial_reconstruction", "gt_adv_split_u", "gt_adv_split_v",
    "gt_adv_split_w", "gt_adv_upstream_2dh", "gt_do_advection",
    "gt_do_advection_3d", "gt_eos_linear", "gt_ip_blumberg_mellor",
    "gt_ip_shchepetkin_mcwilliams", "gt_do_internal_pressure",
    "gt_update_3d_halo", "gt_update_2d_halo", "gt_wait_halo",
    "gt_exchange_tiles", "gt_budget", "gt_malloc", "gt_free",
    "gt_memcpy_h2d", "gt_memcpy_d2h", "gt_adv3d_batch_f32", "gt_adv3d_batch_f64",
    "gt_set_tuning", "gt_set_adjoint_mode", "gt_budget_f32", "gt_adv3d_adjoint",
    "gt_adv3d_adjoint_f32",
    "gt_adv3d_adjoint_seq", "gt_adv3d_adjoint_seq_f32", "gt_ip_shchepetkin_mcwilliams_adjoint",
    "gt_ip_blumberg_mellor_adjoint", "gt_eos_linear_adjoint", "gt_autotune_batch",
    "gt_do_eqstate", "gt_rho_from_theta_unesco80", "gt_rho_from_theta", "gt_eosall_from_theta",
    "gt_start_macro", "gt_set_coordinates", "gt_coordinates", "gt_ww_momentum_3d",
    "gt_salinity_diffusion", "gt_temperature_diffusion", "gt_tridiagonal_batch",
    "gt_uv_advect_3d", "gt_restart_write_host", "gt_restart_read_host", "gt_write_restart",
    "gt_read_restart",
]


class BatchDomain(C.Structure):
    """gt_batch_domain of include/getm_transport.h"""
    _fields_ = [("ctx", C.c_void_p), ("f_in", C.c_void_p), ("f_out", C.c_void_p),
                ("uu", C.c_void_p), ("vv", C.c_void_p), ("ww", C.c_void_p),
                ("hu", C.c_void_p), ("hv", C.c_void_p), ("ho", C.c_void_p),
                ("hires", C.c_void_p)]


class RestartField(C.Structure):
    """gt_restart_field of include/getm_transport.h"""
    _fields_ = [("name", C.c_char_p), ("ndim", C.c_int), ("data", C.c_void_p)]


class GtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{GT_ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def library_path() -> Path:
    return _build.BUILD / _build.LIB_NAME


def load_library():
    """Load libgetm_transport.so (built in-tree by code_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    p = library_path()
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `python code_amd/build.py` "
                          "(the transport step has no CPU fallback)")
    lib = C.CDLL(str(p))
    vp, d, i, l = C.c_void_p, C.c_double, C.c_int, C.c_long
    sig = {
        "gt_last_error": (C.c_char_p, []),
        "gt_version": (i, []),
        "gt_create": (i, [C.POINTER(vp), i, i, i, i, i, d, d, vp, vp, vp, vp]),
        "gt_destroy": (i, [vp]),
        "gt_set_stream": (i, [vp, vp]),
        "gt_synchronize": (i, [vp]),
        "gt_set_adv_ver_iterations": (i, [vp, i]),
        "gt_warnings": (C.c_longlong, [vp]),
        "gt_adv_interfacial_reconstruction": (i, [vp, i, l, vp, vp, vp, vp, vp]),
        "gt_adv_split_u": (i, [vp, d, vp, vp, vp, vp, vp, vp, d, i, d, i]),
        "gt_adv_split_v": (i, [vp, d, vp, vp, vp, vp, vp, vp, d, i, d, i]),
        "gt_adv_split_w": (i, [vp, d, vp, vp, vp, vp, vp, d, i, i, i]),
        "gt_adv_upstream_2dh": (i, [vp, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, i]),
        "gt_do_advection": (i, [vp, d, vp, vp, vp, vp, vp, vp, vp, i, i, d, i, vp, vp]),
        "gt_do_advection_3d": (i, [vp, d, vp, vp, vp, vp, vp, vp, vp, vp, i, i, i, d, i,
                                   vp, vp]),
        "gt_eos_linear": (i, [vp, vp, vp, vp, vp, d, d, d, d]),
        "gt_do_eqstate": (i, [vp, i, i, vp, vp, vp, vp, vp, vp, vp, d, d, d, d]),
        "gt_start_macro": (i, [vp, i, d] + [vp] * 15),
        "gt_set_coordinates": (i, [vp, i, d, d, d, i, d]),
        "gt_coordinates": (i, [vp, d, d, d] + [vp] * 16),
        "gt_ww_momentum_3d": (i, [vp, d, vp, vp, vp, vp, vp]),
        "gt_salinity_diffusion": (i, [vp, d, d, vp, vp, vp]),
        "gt_uv_advect_3d": (i, [vp, d] + [vp] * 7 + [i, i, i, vp, vp]),
        "gt_restart_write_host": (i, [C.c_char_p, i, i, i, i, i, i, i, i, vp]),
        "gt_restart_read_host": (i, [C.c_char_p, i, i, i, i, i, i, vp, i, vp]),
        "gt_write_restart": (i, [vp, C.c_char_p, i, i, i, vp]),
        "gt_read_restart": (i, [vp, C.c_char_p, i, vp, i, vp]),
        "gt_temperature_diffusion": (i, [vp, d, d] + [vp] * 9 + [d, d, vp, vp]),
        "gt_tridiagonal_batch": (i, [vp, C.c_long, i, i, vp, vp, vp, vp, vp]),
        "gt_rho_from_theta_unesco80": (i, [vp, C.c_long, vp, vp, vp, i]),
        "gt_rho_from_theta": (i, [vp, C.c_long, vp, vp, vp, vp, vp]),
        "gt_eosall_from_theta": (i, [vp, C.c_long, vp, vp, vp, vp, vp]),
        "gt_ip_blumberg_mellor": (i, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gt_ip_shchepetkin_mcwilliams": (i, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gt_do_internal_pressure": (i, [vp, i, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gt_update_3d_halo": (i, [vp, vp, vp, i, i]),
        "gt_update_2d_halo": (i, [vp, vp, vp, i, i]),
        "gt_wait_halo": (i, [vp, i]),
        "gt_exchange_tiles": (i, [i, i, vp, vp, i, vp, vp]),
        "gt_budget": (i, [vp, vp, vp, C.POINTER(d), C.POINTER(d)]),
        "gt_malloc": (i, [C.POINTER(vp), C.c_size_t]),
        "gt_free": (i, [vp]),
        "gt_memcpy_h2d": (i, [vp, vp, C.c_size_t]),
        "gt_memcpy_d2h": (i, [vp, vp, C.c_size_t]),
        "gt_adv3d_batch_f32": (i, [i, C.POINTER(BatchDomain), i, d, i, i, d, i, vp]),
        "gt_adv3d_batch_f64": (i, [i, C.POINTER(BatchDomain), i, d, i, i, d, i, vp]),
        "gt_set_tuning": (i, [vp, i, i, i]),
        "gt_set_adjoint_mode": (i, [vp, i]),
        "gt_budget_f32": (i, [vp, vp, vp, C.POINTER(d), C.POINTER(d)]),
        "gt_adv3d_adjoint": (i, [vp, i, i, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i, i,
                                 d, i]),
        "gt_adv3d_adjoint_f32": (i, [vp, i, i, i, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                     i, i, d, i]),
        "gt_adv3d_adjoint_seq": (i, [vp, i, i, d, vp, vp, vp, vp] + [C.POINTER(vp)] * 6 +
                                 [i, i, d, i]),
        "gt_adv3d_adjoint_seq_f32": (i, [vp, i, i, i, d, vp, vp, vp, vp] + [C.POINTER(vp)] * 6 +
                                     [i, i, d, i]),
        "gt_ip_blumberg_mellor_adjoint": (i, [vp, vp, vp, vp, vp, vp, vp]),
        "gt_ip_shchepetkin_mcwilliams_adjoint": (i, [vp] * 10),
        "gt_eos_linear_adjoint": (i, [vp, vp, vp, vp, d, d]),
        "gt_autotune_batch": (i, [i, C.POINTER(BatchDomain), i, d, i, i, d, i, i, i, vp,
                                  C.POINTER(i), C.POINTER(i), C.POINTER(C.c_float),
                                  C.POINTER(C.c_float)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(t):
    """device pointer of a tensor (or None)"""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def check(rc):
    if rc != 0:
        raise GtError(rc, load_library().gt_last_error().decode())


def _restart_fields(fields, ptr):
    """dict name -> array (kmax+1, NY, NX) or (NY, NX) -> gt_restart_field[]"""
    arr = (RestartField * max(1, len(fields)))()
    keep = []
    for q, (name, a) in enumerate(fields.items()):
        nm = name.encode()
        keep.append(nm)
        arr[q] = RestartField(nm, 3 if a.ndim == 3 else 2, ptr(a))
    return arr, keep


def restart_write_host(path, imin, imax, jmin, jmax, kmax, fields, loop=0, with_halos=False):
    """gt_restart_write_host on C-contiguous float64 numpy arrays"""
    arr, keep = _restart_fields(fields, lambda a: a.ctypes.data)
    check(load_library().gt_restart_write_host(str(path).encode(), imin, imax, jmin, jmax, kmax,
                                               int(with_halos), int(loop), len(fields), arr))


def restart_read_host(path, imin, imax, jmin, jmax, kmax, fields, with_halos=False):
    """gt_restart_read_host into numpy arrays; returns loop"""
    arr, keep = _restart_fields(fields, lambda a: a.ctypes.data)
    loop = C.c_int()
    check(load_library().gt_restart_read_host(str(path).encode(), imin, imax, jmin, jmax, kmax,
                                              int(with_halos), C.byref(loop), len(fields), arr))
    return loop.value


class Context:
    """One domain on the device (gt_create).

    ``az``, ``au``, ``av``, ``ax`` are (NY, NX) int32 numpy arrays; ``au``,
    ``av``, ``ax`` may be None (derived from az as domain.F90 does).
    """

    def __init__(self, imax, jmax, kmax, dx, dy, az, au=None, av=None, ax=None,
                 imin=1, jmin=1, stream=None):
        lib = load_library()
        self.lib = lib
        self.imin, self.imax, self.jmin, self.jmax, self.kmax = imin, imax, jmin, jmax, kmax
        self.NX = imax - imin + 5
        self.NY = jmax - jmin + 5
        self.dx, self.dy = float(dx), float(dy)

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.int32)
            assert a.shape == (self.NY, self.NX), (a.shape, (self.NY, self.NX))
            return a
        self._masks = [arr(m) for m in (az, au, av, ax)]
        ptrs = [None if m is None else m.ctypes.data_as(C.c_void_p) for m in self._masks]
        h = C.c_void_p()
        check(lib.gt_create(C.byref(h), imin, imax, jmin, jmax, kmax, self.dx, self.dy,
                            *ptrs))
        self.h = h
        if stream is not None:
            self.set_stream(stream)

    @classmethod
    def from_domain(cls, dom, **kw):
        return cls(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, dom.az, dom.au, dom.av,
                   dom.ax, imin=dom.imin, jmin=dom.jmin, **kw)

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.gt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream (or raw handle int)"""
        handle = getattr(stream, "cuda_stream", stream)
        check(self.lib.gt_set_stream(self.h, C.c_void_p(handle)))
        self.stream = stream

    def synchronize(self):
        check(self.lib.gt_synchronize(self.h))

    def set_adv_ver_iterations(self, n):
        check(self.lib.gt_set_adv_ver_iterations(self.h, int(n)))

    def warnings(self):
        return int(self.lib.gt_warnings(self.h))

    # ---- advection --------------------------------------------------------
    def adv_interfacial_reconstruction(self, scheme, cfl, fuu, fu, fd, out):
        check(self.lib.gt_adv_interfacial_reconstruction(
            self.h, scheme, cfl.numel(), _ptr(cfl), _ptr(fuu), _ptr(fu), _ptr(fd), _ptr(out)))

    def adv_split_u(self, dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH, tag=H_TAG):
        check(self.lib.gt_adv_split_u(self.h, dt, _ptr(f), _ptr(fi), _ptr(Di), _ptr(adv),
                                      _ptr(U), _ptr(DU), splitfac, scheme, AH, tag))

    def adv_split_v(self, dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH, tag=H_TAG):
        check(self.lib.gt_adv_split_v(self.h, dt, _ptr(f), _ptr(fi), _ptr(Di), _ptr(adv),
                                      _ptr(V), _ptr(DV), splitfac, scheme, AH, tag))

    def adv_split_w(self, dt, f, fi, hi, adv, ww, splitfac, scheme, tag, itersmax):
        check(self.lib.gt_adv_split_w(self.h, dt, _ptr(f), _ptr(fi), _ptr(hi), _ptr(adv),
                                      _ptr(ww), splitfac, scheme, tag, itersmax))

    def adv_upstream_2dh(self, dt, f, fi, Di, adv, U, V, Dn, DU, DV, AH, tag=H_TAG):
        check(self.lib.gt_adv_upstream_2dh(self.h, dt, _ptr(f), _ptr(fi), _ptr(Di),
                                           _ptr(adv), _ptr(U), _ptr(V), _ptr(Dn), _ptr(DU),
                                           _ptr(DV), AH, tag))

    def do_advection(self, dt, f, U, V, DU, DV, Do, Dn, split, scheme, AH, tag,
                     Dires=None, advres=None):
        check(self.lib.gt_do_advection(self.h, dt, _ptr(f), _ptr(U), _ptr(V), _ptr(DU),
                                       _ptr(DV), _ptr(Do), _ptr(Dn), split, scheme, AH, tag,
                                       _ptr(Dires), _ptr(advres)))

    def do_advection_3d(self, dt, f, uu, vv, ww, hu, hv, ho, hn, split, hscheme, vscheme,
                        AH, tag, hires=None, advres=None):
        check(self.lib.gt_do_advection_3d(self.h, dt, _ptr(f), _ptr(uu), _ptr(vv), _ptr(ww),
                                          _ptr(hu), _ptr(hv), _ptr(ho), _ptr(hn), split,
                                          hscheme, vscheme, AH, tag, _ptr(hires),
                                          _ptr(advres)))

    # ---- internal pressure --------------------------------------------------
    def eos_linear(self, T, S, rho, buoy, T0, S0, dtr0, dsr0):
        check(self.lib.gt_eos_linear(self.h, _ptr(T), _ptr(S), _ptr(rho), _ptr(buoy),
                                     T0, S0, dtr0, dsr0))

    def do_eqstate(self, method, T, S, hn, rho, buoy, alpha=None, beta=None, T0=10.0,
                   S0=33.75, dtr0=-0.17, dsr0=0.78, nonnegsalt=False):
        """do_eqstate (eqstate.F90:94-271); the defaults are the reference's
        namelist defaults, FP32 literals (rounded here)."""
        f32 = lambda v: float(np.float32(v))
        check(self.lib.gt_do_eqstate(self.h, method, int(nonnegsalt), _ptr(T), _ptr(S),
                                     _ptr(hn), _ptr(rho), _ptr(buoy), _ptr(alpha), _ptr(beta),
                                     f32(T0), f32(S0), f32(dtr0), f32(dsr0)))

    # ---- input producers (m3d.F90:446-548) ----------------------------------
    def start_macro(self, M, min_depth, H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn, Dn,
                    Dun, Dvn, Uint, Vint):
        check(self.lib.gt_start_macro(self.h, int(M), min_depth,
                                      *(_ptr(a) for a in (H, HU, HV, z, sseo, ssen, ssuo, ssun,
                                                          ssvo, ssvn, Dn, Dun, Dvn, Uint, Vint))))

    def set_coordinates(self, vert_cord, ddu=0.0, ddl=0.0, d_gamma=20.0, gamma_surf=True,
                        min_depth=0.5):
        check(self.lib.gt_set_coordinates(self.h, int(vert_cord), ddu, ddl, d_gamma,
                                          int(gamma_surf), min_depth))

    def coordinates(self, dt, cord_relax, maxdepth, H, HU, HV, sseo, ssen, ssuo, ssun, ssvo,
                    ssvn, ho, hn, huo, hun, hvo, hvn, zc=None):
        check(self.lib.gt_coordinates(self.h, dt, cord_relax, maxdepth,
                                      *(_ptr(a) for a in (H, HU, HV, sseo, ssen, ssuo, ssun,
                                                          ssvo, ssvn, ho, hn, huo, hun, hvo, hvn,
                                                          zc))))

    def ww_momentum_3d(self, dt, uu, vv, ho, hn, ww):
        check(self.lib.gt_ww_momentum_3d(self.h, dt, _ptr(uu), _ptr(vv), _ptr(ho), _ptr(hn),
                                         _ptr(ww)))

    # ---- restart fields (the NetCDF hotstart file) ----------------------------
    def write_restart(self, path, fields, loop=0, with_halos=False):
        arr, keep = _restart_fields(fields, lambda t: t.data_ptr())
        check(self.lib.gt_write_restart(self.h, str(path).encode(), int(with_halos), int(loop),
                                        len(fields), arr))

    def read_restart(self, path, fields, with_halos=False):
        arr, keep = _restart_fields(fields, lambda t: t.data_ptr())
        loop = C.c_int()
        check(self.lib.gt_read_restart(self.h, str(path).encode(), int(with_halos),
                                       C.byref(loop), len(fields), arr))
        return loop.value

    def uv_advect_3d(self, dt, uu, vv, ww, ho, hn, hun, hvn, split, hscheme, vscheme, uuEx,
                     vvEx):
        check(self.lib.gt_uv_advect_3d(self.h, dt, *(_ptr(a) for a in (uu, vv, ww, ho, hn, hun,
                                                                       hvn)),
                                       split, hscheme, vscheme, _ptr(uuEx), _ptr(vvEx)))

    # ---- vertical diffusion (temperature.F90:463-551, salinity.F90:346-393) ----
    def salinity_diffusion(self, dt, cnpar, hn, nuh, S):
        check(self.lib.gt_salinity_diffusion(self.h, dt, cnpar, _ptr(hn), _ptr(nuh), _ptr(S)))

    def temperature_diffusion(self, dt, cnpar, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T,
                              swr_bot_refl_frac=-1.0, swr_min_bot_frac=0.01):
        f32 = lambda v: float(np.float32(v))
        check(self.lib.gt_temperature_diffusion(
            self.h, dt, cnpar, *(_ptr(x) for x in (hn, nuh, swr, shf, A, g1, g2, dry_z, S)),
            swr_bot_refl_frac, f32(swr_min_bot_frac), _ptr(rad), _ptr(T)))

    def tridiagonal_batch(self, fi, lt, au, bu, cu, du, value):
        """systems along the last axis: arrays of shape (lt+1, n)"""
        check(self.lib.gt_tridiagonal_batch(self.h, au.shape[-1], fi, lt, _ptr(au), _ptr(bu),
                                            _ptr(cu), _ptr(du), _ptr(value)))

    def rho_from_theta_unesco80(self, T, S, rho, nonnegsalt=False):
        check(self.lib.gt_rho_from_theta_unesco80(self.h, T.numel(), _ptr(T), _ptr(S),
                                                  _ptr(rho), int(nonnegsalt)))

    def rho_from_theta(self, s, th, p, dens0, densp=None):
        check(self.lib.gt_rho_from_theta(self.h, s.numel(), _ptr(s), _ptr(th), _ptr(p),
                                         _ptr(dens0), _ptr(densp)))

    def eosall_from_theta(self, s, th, p, rho_s, rho_th):
        check(self.lib.gt_eosall_from_theta(self.h, s.numel(), _ptr(s), _ptr(th), _ptr(p),
                                            _ptr(rho_s), _ptr(rho_th)))

    def ip_blumberg_mellor(self, H, hn, hun, hvn, buoy, zz, idpdx, idpdy):
        check(self.lib.gt_ip_blumberg_mellor(self.h, _ptr(H), _ptr(hn), _ptr(hun), _ptr(hvn),
                                             _ptr(buoy), _ptr(zz), _ptr(idpdx), _ptr(idpdy)))

    def ip_shchepetkin_mcwilliams(self, H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy):
        check(self.lib.gt_ip_shchepetkin_mcwilliams(
            self.h, _ptr(H), _ptr(hn), _ptr(hun), _ptr(hvn), _ptr(buoy), _ptr(sseo), _ptr(zz),
            _ptr(idpdx), _ptr(idpdy)))

    def do_internal_pressure(self, ip_method, H, hn, hun, hvn, buoy, sseo, idpdx, idpdy):
        check(self.lib.gt_do_internal_pressure(
            self.h, ip_method, _ptr(H), _ptr(hn), _ptr(hun), _ptr(hvn), _ptr(buoy), _ptr(sseo),
            _ptr(idpdx), _ptr(idpdy)))

    # ---- ghost cells -----------------------------------------------------------
    def update_3d_halo(self, f1, f2, tag=H_TAG, mirror=-1):
        check(self.lib.gt_update_3d_halo(self.h, _ptr(f1), _ptr(f2), tag, mirror))

    def update_2d_halo(self, f1, f2, tag=H_TAG, mirror=-1):
        check(self.lib.gt_update_2d_halo(self.h, _ptr(f1), _ptr(f2), tag, mirror))

    def wait_halo(self, tag=H_TAG):
        check(self.lib.gt_wait_halo(self.h, tag))

    # ---- budgets ---------------------------------------------------------------
    def budget(self, h, f=None):
        """(mass, volume); f may be an FP64 or FP32 field"""
        import torch
        m, v = C.c_double(), C.c_double()
        fn = self.lib.gt_budget_f32 if (f is not None and f.dtype == torch.float32) \
            else self.lib.gt_budget
        check(fn(self.h, _ptr(h), _ptr(f), C.byref(m), C.byref(v)))
        return m.value, v.value

    # ---- reverse mode -------------------------------------------------------------
    def adv3d_adjoint(self, nsteps, ckpt_every, dt, f0, lam_final, lam0, uu, vv, ww, hu, hv, ho,
                      hscheme, vscheme, AH=-1.0, tag=H_TAG, f_final=None):
        check(self.lib.gt_adv3d_adjoint(self.h, nsteps, ckpt_every, dt, _ptr(f0), _ptr(f_final),
                                        _ptr(lam_final), _ptr(lam0), _ptr(uu), _ptr(vv),
                                        _ptr(ww), _ptr(hu), _ptr(hv), _ptr(ho), hscheme,
                                        vscheme, AH, tag))

    def adv3d_adjoint_f32(self, ntracers, nsteps, ckpt_every, dt, f0, lam_final, lam0, uu, vv,
                          ww, hu, hv, ho, hscheme, vscheme, AH=-1.0, tag=H_TAG, f_final=None):
        """gt_adv3d_adjoint_f32: f0, lam_final, lam0 (and f_final) are FP32
        tensors of shape (ntracers, kmax+1, NY, NX)"""
        import torch
        n3 = (self.kmax + 1) * self.NY * self.NX
        for t in (f0, lam_final, lam0, f_final):
            if t is not None and (t.dtype != torch.float32 or t.numel() != ntracers * n3
                                  or not t.is_contiguous()):
                raise ValueError("adv3d_adjoint_f32: tracer sets must be contiguous FP32 "
                                 "tensors of ntracers x (kmax+1) x NY x NX")
        check(self.lib.gt_adv3d_adjoint_f32(self.h, ntracers, nsteps, ckpt_every, dt, _ptr(f0),
                                            _ptr(f_final), _ptr(lam_final), _ptr(lam0), _ptr(uu),
                                            _ptr(vv), _ptr(ww), _ptr(hu), _ptr(hv), _ptr(ho),
                                            hscheme, vscheme, AH, tag))

    def adv3d_adjoint_seq(self, ntracers, ckpt_every, dt, f0, lam_final, lam0, steps, hscheme,
                          vscheme, AH=-1.0, tag=H_TAG, f_final=None):
        """gt_adv3d_adjoint_seq (FP64 tensors, ntracers = 1) or _seq_f32 (FP32,
        shape (ntracers, kmax+1, NY, NX)); steps: one (uu, vv, ww, hu, hv, ho)
        tuple of device tensors per step"""
        import torch
        n = len(steps)
        arrs = []
        for q in range(6):
            a = (C.c_void_p * n)(*[steps[t][q].data_ptr() for t in range(n)])
            arrs.append(a)
        ptrs = [C.cast(a, C.POINTER(C.c_void_p)) for a in arrs]
        self._seq_keep = arrs  # alive until the call returns
        if f0.dtype == torch.float32:
            check(self.lib.gt_adv3d_adjoint_seq_f32(self.h, ntracers, n, ckpt_every, dt, _ptr(f0),
                                                    _ptr(f_final), _ptr(lam_final), _ptr(lam0),
                                                    *ptrs, hscheme, vscheme, AH, tag))
        else:
            if ntracers != 1:
                raise ValueError("adv3d_adjoint_seq: FP64 takes one tracer")
            check(self.lib.gt_adv3d_adjoint_seq(self.h, n, ckpt_every, dt, _ptr(f0), _ptr(f_final),
                                                _ptr(lam_final), _ptr(lam0), *ptrs, hscheme,
                                                vscheme, AH, tag))

    def ip_blumberg_mellor_adjoint(self, hun, hvn, zz, idpdx_bar, idpdy_bar, buoy_bar):
        check(self.lib.gt_ip_blumberg_mellor_adjoint(self.h, _ptr(hun), _ptr(hvn), _ptr(zz),
                                                     _ptr(idpdx_bar), _ptr(idpdy_bar),
                                                     _ptr(buoy_bar)))

    def ip_shchepetkin_mcwilliams_adjoint(self, hn, hun, hvn, buoy, sseo, zz, idpdx_bar,
                                          idpdy_bar, buoy_bar):
        check(self.lib.gt_ip_shchepetkin_mcwilliams_adjoint(
            self.h, _ptr(hn), _ptr(hun), _ptr(hvn), _ptr(buoy), _ptr(sseo), _ptr(zz),
            _ptr(idpdx_bar), _ptr(idpdy_bar), _ptr(buoy_bar)))

    def eos_linear_adjoint(self, buoy_bar, T_bar, S_bar, dtr0, dsr0):
        check(self.lib.gt_eos_linear_adjoint(self.h, _ptr(buoy_bar), _ptr(T_bar), _ptr(S_bar),
                                             dtr0, dsr0))

    def set_tuning(self, fused=1, tile=0, group=3):
        check(self.lib.gt_set_tuning(self.h, int(fused), int(tile), int(group)))

    def set_adjoint_mode(self, keep_states=-1):
        """gt_set_adjoint_mode: 1 keeps each segment's sub-step states, 0
        recomputes them per reverse step, -1 decides by free memory"""
        check(self.lib.gt_set_adjoint_mode(self.h, int(keep_states)))


def exchange_tiles(px, py, ni, nj, nk, tiles, stream=None):
    """gt_exchange_tiles on a list of device tensors"""
    lib = load_library()
    n = px * py
    ni_a = (C.c_int * n)(*ni)
    nj_a = (C.c_int * n)(*nj)
    tp = (C.c_void_p * n)(*[t.data_ptr() for t in tiles])
    handle = getattr(stream, "cuda_stream", stream)
    check(lib.gt_exchange_tiles(px, py, ni_a, nj_a, nk, tp,
                                C.c_void_p(handle) if handle else None))


def _batch_array(domains):
    n = len(domains)
    arr = (BatchDomain * n)()
    for q, d in enumerate(domains):
        arr[q] = BatchDomain(d["ctx"].h.value, d["f_in"].data_ptr(), d["f_out"].data_ptr(),
                             *(d[k].data_ptr() for k in ("uu", "vv", "ww", "hu", "hv", "ho")),
                             d["hires"].data_ptr() if d.get("hires") is not None else None)
    return arr


# strip heights of the HALFSPLIT strip kernels; a tuning index t < 3 gives
# ROWS[t] rows to both kernels, 3 <= t < 12 gives ROWS[(t-3)%3] to the first
# (u/2 v/2) and ROWS[(t-3)//3] to the second (w v/2 u/2)
ROWS = (16, 32, 8)


def strip_rows(tile):
    """(rows of the first kernel, rows of the second) for a tuning index"""
    if 3 <= tile < 12:
        return ROWS[(tile - 3) % 3], ROWS[(tile - 3) // 3]
    r = ROWS[tile if 0 <= tile < 3 else 0]
    return r, r


def tile_for_rows(rows_uv, rows_wvu=None):
    """the tuning index of a strip-height pair (8, 16 or 32 rows each)"""
    rows_wvu = rows_uv if rows_wvu is None else rows_wvu
    if rows_uv not in ROWS or rows_wvu not in ROWS:
        raise ValueError(f"rows per strip must be one of {sorted(ROWS)}")
    return 3 + ROWS.index(rows_uv) + 3 * ROWS.index(rows_wvu)


def autotune_batch(domains, ntracers, dt, hscheme, vscheme, AH=-1.0, tag=H_TAG, reps=3,
                   stream=None):
    """gt_autotune_batch: returns (tile, group, best_ms, {(tile, group): ms}); tile
    is a strip-height index (strip_rows), group the tracers per wavefront"""
    import torch
    lib = load_library()
    arr = _batch_array(domains)
    f32 = domains[0]["f_in"].dtype == torch.float32
    bt, bg, bm = C.c_int(), C.c_int(), C.c_float()
    table = (C.c_float * 36)()
    handle = getattr(stream, "cuda_stream", stream)
    check(lib.gt_autotune_batch(len(domains), arr, ntracers, dt, hscheme, vscheme, AH, tag,
                                int(f32), reps, C.c_void_p(handle) if handle else None,
                                C.byref(bt), C.byref(bg), C.byref(bm), table))
    pairs = [(t, g) for t in range(3, 12) for g in ((1, 3, 6, 9) if f32 else (1, 2))]
    return bt.value, bg.value, bm.value, {p: table[i] for i, p in enumerate(pairs)}


def adv3d_batch(domains, ntracers, dt, hscheme, vscheme, AH=-1.0, tag=H_TAG, stream=None):
    """gt_adv3d_batch_f32/_f64.  domains: list of dicts with keys ctx, f_in,
    f_out (tensors of shape (ntracers, kmax+1, NY, NX)), uu, vv, ww, hu, hv,
    ho and optional hires.  The dtype of f_in picks the FP32 or FP64 kernel."""
    import torch
    lib = load_library()
    n = len(domains)
    arr = _batch_array(domains)
    f32 = domains[0]["f_in"].dtype == torch.float32
    fn = lib.gt_adv3d_batch_f32 if f32 else lib.gt_adv3d_batch_f64
    handle = getattr(stream, "cuda_stream", stream)
    check(fn(n, arr, ntracers, dt, hscheme, vscheme, AH, tag,
             C.c_void_p(handle) if handle else None))
