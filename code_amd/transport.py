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
HU_TAG, HV_TAG, DU_TAG, DV_TAG, Z_TAG = 11, 12, 21, 22, 30  # halo_zones.F90:25-27
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
    "gt_read_restart", "gt_set_substr_ini_press", "gt_reserve", "gt_budget_batch_f32",
    "gt_budget_batch_f32_noalloc", "gt_adv_split_u_m", "gt_adv_split_v_m", "gt_adv_split_w_m",
    "gt_adv_upstream_2dh_m", "gt_eos_internal_pressure", "gt_eqstate_internal_pressure",
    "gt_get_masks",
]


class BatchDomain(C.Structure):
    """gt_batch_domain of include/getm_transport.h"""
    _fields_ = [("ctx", C.c_void_p), ("f_in", C.c_void_p), ("f_out", C.c_void_p),
                ("uu", C.c_void_p), ("vv", C.c_void_p), ("ww", C.c_void_p),
                ("hu", C.c_void_p), ("hv", C.c_void_p), ("ho", C.c_void_p),
                ("hires", C.c_void_p), ("budget", C.c_void_p)]


class RestartField(C.Structure):
    """gt_restart_field of include/getm_transport.h"""
    _fields_ = [("name", C.c_char_p), ("ndim", C.c_int), ("data", C.c_void_p)]


class RestartTime(C.Structure):
    """gt_restart_time: the scalars of create_restart_ncdf.F90:94-104"""
    _fields_ = [("loop", C.c_int), ("julianday", C.c_int), ("secondsofday", C.c_int),
                ("timestep", C.c_double)]

    def as_dict(self):
        return dict(loop=self.loop, julianday=self.julianday, secondsofday=self.secondsofday,
                    timestep=self.timestep)


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
        "gt_get_masks": (i, [vp, vp, vp, vp, vp]),
        "gt_set_stream": (i, [vp, vp]),
        "gt_synchronize": (i, [vp]),
        "gt_set_adv_ver_iterations": (i, [vp, i]),
        "gt_warnings": (C.c_longlong, [vp]),
        "gt_adv_interfacial_reconstruction": (i, [vp, i, l, vp, vp, vp, vp, vp]),
        "gt_adv_split_u": (i, [vp, d, vp, vp, vp, vp, vp, vp, d, i, d, i]),
        "gt_adv_split_v": (i, [vp, d, vp, vp, vp, vp, vp, vp, d, i, d, i]),
        "gt_adv_split_w": (i, [vp, d, vp, vp, vp, vp, vp, d, i, i, i]),
        "gt_adv_upstream_2dh": (i, [vp, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, i]),
        "gt_adv_split_u_m": (i, [vp, d, vp, vp, vp, vp, vp, vp, d, i, d, vp, vp]),
        "gt_adv_split_v_m": (i, [vp, d, vp, vp, vp, vp, vp, vp, d, i, d, vp, vp]),
        "gt_adv_split_w_m": (i, [vp, d, vp, vp, vp, vp, vp, d, i, i, vp, i]),
        "gt_adv_upstream_2dh_m": (i, [vp, d, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp]),
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
        "gt_restart_write_host": (i, [C.c_char_p, i, i, i, i, i, i, C.POINTER(RestartTime), i,
                                      vp]),
        "gt_restart_read_host": (i, [C.c_char_p, i, i, i, i, i, i, C.POINTER(RestartTime), i,
                                     vp]),
        "gt_write_restart": (i, [vp, C.c_char_p, i, C.POINTER(RestartTime), i, vp]),
        "gt_read_restart": (i, [vp, C.c_char_p, i, C.POINTER(RestartTime), i, vp]),
        "gt_temperature_diffusion": (i, [vp, d, d] + [vp] * 9 + [d, d, vp, vp]),
        "gt_tridiagonal_batch": (i, [vp, C.c_long, i, i, vp, vp, vp, vp, vp]),
        "gt_rho_from_theta_unesco80": (i, [vp, C.c_long, vp, vp, vp, i]),
        "gt_rho_from_theta": (i, [vp, C.c_long, vp, vp, vp, vp, vp]),
        "gt_eosall_from_theta": (i, [vp, C.c_long, vp, vp, vp, vp, vp]),
        "gt_ip_blumberg_mellor": (i, [vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gt_ip_shchepetkin_mcwilliams": (i, [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gt_do_internal_pressure": (i, [vp, i, vp, vp, vp, vp, vp, vp, vp, vp]),
        "gt_eos_internal_pressure": (i, [vp, i, vp, vp, d, d, d, d] + [vp] * 11),
        "gt_eqstate_internal_pressure": (i, [vp, i, i, i, vp, vp, vp, d, d, d, d] + [vp] * 11),
        "gt_set_substr_ini_press": (i, [vp, i]),
        "gt_reserve": (i, [vp, i, i]),
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
        "gt_budget_batch_f32": (i, [vp, i, vp, vp, vp]),
        "gt_budget_batch_f32_noalloc": (i, [vp, i, vp, vp, vp]),
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
    """device pointer of a tensor (or None); callers validate first"""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _dev(t, name, dtype=None, shape=None, numel=None, optional=False):
    """validated device pointer: a contiguous CUDA tensor of ``dtype`` (FP64
    by default) and ``shape`` (or ``numel`` elements).  Raises ValueError
    instead of handing the library an address it would read out of bounds."""
    import torch
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name}: a device tensor is required")
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: must be a GPU tensor (got {t.device})")
    want = torch.float64 if dtype is None else dtype
    if t.dtype != want:
        raise ValueError(f"{name}: dtype must be {want} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape must be {tuple(shape)} (got {tuple(t.shape)})")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: must hold {numel} elements (got {t.numel()})")
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


def _restart_time(loop, julianday, secondsofday, timestep):
    return RestartTime(int(loop), int(julianday), int(secondsofday), float(timestep))


def restart_write_host(path, imin, imax, jmin, jmax, kmax, fields, loop=0, with_halos=False,
                       julianday=0, secondsofday=0, timestep=0.0):
    """gt_restart_write_host on C-contiguous float64 numpy arrays"""
    arr, keep = _restart_fields(fields, lambda a: a.ctypes.data)
    t = _restart_time(loop, julianday, secondsofday, timestep)
    check(load_library().gt_restart_write_host(str(path).encode(), imin, imax, jmin, jmax, kmax,
                                               int(with_halos), C.byref(t), len(fields), arr))


def restart_read_host(path, imin, imax, jmin, jmax, kmax, fields, with_halos=False):
    """gt_restart_read_host into numpy arrays; returns the scalars as a dict
    (loop, julianday, secondsofday, timestep)"""
    arr, keep = _restart_fields(fields, lambda a: a.ctypes.data)
    t = RestartTime()
    check(load_library().gt_restart_read_host(str(path).encode(), imin, imax, jmin, jmax, kmax,
                                              int(with_halos), C.byref(t), len(fields), arr))
    return t.as_dict()


class Context:
    """One domain on the device (gt_create).

    ``az``, ``au``, ``av``, ``ax`` are (NY, NX) int32 numpy arrays; ``au``,
    ``av``, ``ax`` may be None (derived from az as domain.F90 does).  Every
    method checks its tensors (GPU, dtype, contiguous, shape) before the
    library sees them and raises ValueError otherwise.
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
            if a.shape != (self.NY, self.NX):
                raise ValueError(f"mask shape {a.shape} != {(self.NY, self.NX)}")
            return a
        self._masks = [arr(m) for m in (az, au, av, ax)]
        if self._masks[0] is None:
            raise ValueError("az is required")
        ptrs = [None if m is None else m.ctypes.data_as(C.c_void_p) for m in self._masks]
        h = C.c_void_p()
        check(lib.gt_create(C.byref(h), imin, imax, jmin, jmax, kmax, self.dx, self.dy,
                            *ptrs))
        self.h = h
        if stream is not None:
            self.set_stream(stream)

    def masks(self):
        """(az, au, av, ax) as the context holds them (gt_get_masks)"""
        out = [np.zeros((self.NY, self.NX), np.int32) for _ in range(4)]
        check(self.lib.gt_get_masks(self.h, *(m.ctypes.data_as(C.c_void_p) for m in out)))
        return tuple(out)

    @classmethod
    def from_domain(cls, dom, **kw):
        return cls(dom.imax, dom.jmax, dom.kmax, dom.dx, dom.dy, dom.az, dom.au, dom.av,
                   dom.ax, imin=dom.imin, jmin=dom.jmin, **kw)

    @property
    def shape3(self):
        return (self.kmax + 1, self.NY, self.NX)

    @property
    def shape2(self):
        return (self.NY, self.NX)

    def _f3(self, t, name, optional=False, dtype=None):
        return _dev(t, name, dtype=dtype, shape=self.shape3, optional=optional)

    def _f2(self, t, name, optional=False):
        return _dev(t, name, shape=self.shape2, optional=optional)

    def _set(self, t, name, ntracers, dtype, optional=False):
        """a tracer set: ntracers fields back to back"""
        return _dev(t, name, dtype=dtype, numel=ntracers * int(np.prod(self.shape3)),
                    optional=optional)

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

    def reserve(self, ntracers_f32=0, ntracers_f64=0):
        """gt_reserve: size every work field of the transport step now"""
        check(self.lib.gt_reserve(self.h, int(ntracers_f32), int(ntracers_f64)))

    # ---- advection --------------------------------------------------------
    def adv_interfacial_reconstruction(self, scheme, cfl, fuu, fu, fd, out):
        n = cfl.numel() if cfl is not None else 0
        ps = [_dev(t, nm, numel=n) for t, nm in ((cfl, "cfl"), (fuu, "fuu"), (fu, "fu"),
                                                  (fd, "fd"), (out, "out"))]
        check(self.lib.gt_adv_interfacial_reconstruction(self.h, scheme, n, *ps))

    def adv_split_u(self, dt, f, fi, Di, adv, U, DU, splitfac, scheme, AH, tag=H_TAG):
        ps = [self._f2(t, nm) for t, nm in ((f, "f"), (fi, "fi"), (Di, "Di"), (adv, "adv"),
                                            (U, "U"), (DU, "DU"))]
        check(self.lib.gt_adv_split_u(self.h, dt, *ps, splitfac, scheme, AH, tag))

    def adv_split_v(self, dt, f, fi, Di, adv, V, DV, splitfac, scheme, AH, tag=H_TAG):
        ps = [self._f2(t, nm) for t, nm in ((f, "f"), (fi, "fi"), (Di, "Di"), (adv, "adv"),
                                            (V, "V"), (DV, "DV"))]
        check(self.lib.gt_adv_split_v(self.h, dt, *ps, splitfac, scheme, AH, tag))

    def adv_split_w(self, dt, f, fi, hi, adv, ww, splitfac, scheme, tag, itersmax):
        ps = [self._f3(t, nm) for t, nm in ((f, "f"), (fi, "fi"), (hi, "hi"), (adv, "adv"),
                                            (ww, "ww"))]
        check(self.lib.gt_adv_split_w(self.h, dt, *ps, splitfac, scheme, tag, itersmax))

    def adv_upstream_2dh(self, dt, f, fi, Di, adv, U, V, Dn, DU, DV, AH, tag=H_TAG):
        ps = [self._f2(t, nm) for t, nm in ((f, "f"), (fi, "fi"), (Di, "Di"), (adv, "adv"),
                                            (U, "U"), (V, "V"), (Dn, "Dn"))]
        ps += [self._f2(t, nm, optional=True) for t, nm in ((DU, "DU"), (DV, "DV"))]
        check(self.lib.gt_adv_upstream_2dh(self.h, dt, *ps, AH, tag))

    def do_advection(self, dt, f, U, V, DU, DV, Do, Dn, split, scheme, AH, tag,
                     Dires=None, advres=None):
        ps = [self._f2(t, nm) for t, nm in ((f, "f"), (U, "U"), (V, "V"), (DU, "DU"),
                                            (DV, "DV"), (Do, "Do"), (Dn, "Dn"))]
        check(self.lib.gt_do_advection(self.h, dt, *ps, split, scheme, AH, tag,
                                       self._f2(Dires, "Dires", True),
                                       self._f2(advres, "advres", True)))

    def do_advection_3d(self, dt, f, uu, vv, ww, hu, hv, ho, hn, split, hscheme, vscheme,
                        AH, tag, hires=None, advres=None):
        ps = [self._f3(t, nm) for t, nm in ((f, "f"), (uu, "uu"), (vv, "vv"), (ww, "ww"),
                                            (hu, "hu"), (hv, "hv"), (ho, "ho"), (hn, "hn"))]
        check(self.lib.gt_do_advection_3d(self.h, dt, *ps, split, hscheme, vscheme, AH, tag,
                                          self._f3(hires, "hires", True),
                                          self._f3(advres, "advres", True)))

    # ---- internal pressure --------------------------------------------------
    def eos_linear(self, T, S, rho, buoy, T0, S0, dtr0, dsr0):
        check(self.lib.gt_eos_linear(self.h, self._f3(T, "T"), self._f3(S, "S"),
                                     self._f3(rho, "rho", True), self._f3(buoy, "buoy"),
                                     T0, S0, dtr0, dsr0))

    def do_eqstate(self, method, T, S, hn, rho, buoy, alpha=None, beta=None, T0=10.0,
                   S0=33.75, dtr0=-0.17, dsr0=0.78, nonnegsalt=False):
        """do_eqstate (eqstate.F90:94-271); the defaults are the reference's
        namelist defaults, FP32 literals (rounded here)."""
        f32 = lambda v: float(np.float32(v))
        check(self.lib.gt_do_eqstate(self.h, method, int(nonnegsalt), self._f3(T, "T"),
                                     self._f3(S, "S"), self._f3(hn, "hn", True),
                                     self._f3(rho, "rho"), self._f3(buoy, "buoy"),
                                     self._f3(alpha, "alpha", True),
                                     self._f3(beta, "beta", True),
                                     f32(T0), f32(S0), f32(dtr0), f32(dsr0)))

    # ---- input producers (m3d.F90:446-548) ----------------------------------
    def start_macro(self, M, min_depth, H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn, Dn,
                    Dun, Dvn, Uint, Vint):
        names = ("H", "HU", "HV", "z", "sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn", "Dn",
                 "Dun", "Dvn", "Uint", "Vint")
        vals = (H, HU, HV, z, sseo, ssen, ssuo, ssun, ssvo, ssvn, Dn, Dun, Dvn, Uint, Vint)
        check(self.lib.gt_start_macro(self.h, int(M), min_depth,
                                      *(self._f2(a, nm) for a, nm in zip(vals, names))))

    def set_coordinates(self, vert_cord, ddu=0.0, ddl=0.0, d_gamma=20.0, gamma_surf=True,
                        min_depth=0.5):
        check(self.lib.gt_set_coordinates(self.h, int(vert_cord), ddu, ddl, d_gamma,
                                          int(gamma_surf), min_depth))

    def coordinates(self, dt, cord_relax, maxdepth, H, HU, HV, sseo, ssen, ssuo, ssun, ssvo,
                    ssvn, ho, hn, huo, hun, hvo, hvn, zc=None):
        p2 = [self._f2(a, nm, optional=nm != "H") for a, nm in
              zip((H, HU, HV, sseo, ssen, ssuo, ssun, ssvo, ssvn),
                  ("H", "HU", "HV", "sseo", "ssen", "ssuo", "ssun", "ssvo", "ssvn"))]
        p3 = [self._f3(a, nm, optional=nm != "hn") for a, nm in
              zip((ho, hn, huo, hun, hvo, hvn, zc), ("ho", "hn", "huo", "hun", "hvo", "hvn",
                                                     "zc"))]
        check(self.lib.gt_coordinates(self.h, dt, cord_relax, maxdepth, *p2, *p3))

    def ww_momentum_3d(self, dt, uu, vv, ho, hn, ww):
        ps = [self._f3(t, nm) for t, nm in ((uu, "uu"), (vv, "vv"), (ho, "ho"), (hn, "hn"),
                                            (ww, "ww"))]
        check(self.lib.gt_ww_momentum_3d(self.h, dt, *ps))

    # ---- restart fields (the NetCDF hotstart file) ----------------------------
    def _restart_ptr(self, t):
        shape = self.shape3 if t.ndim == 3 else self.shape2
        return _dev(t, "restart field", shape=shape).value

    def write_restart(self, path, fields, loop=0, with_halos=False, julianday=0,
                      secondsofday=0, timestep=0.0):
        arr, keep = _restart_fields(fields, self._restart_ptr)
        t = _restart_time(loop, julianday, secondsofday, timestep)
        check(self.lib.gt_write_restart(self.h, str(path).encode(), int(with_halos), C.byref(t),
                                        len(fields), arr))

    def read_restart(self, path, fields, with_halos=False):
        """returns the scalars as a dict (loop, julianday, secondsofday,
        timestep)"""
        arr, keep = _restart_fields(fields, self._restart_ptr)
        t = RestartTime()
        check(self.lib.gt_read_restart(self.h, str(path).encode(), int(with_halos), C.byref(t),
                                       len(fields), arr))
        return t.as_dict()

    def uv_advect_3d(self, dt, uu, vv, ww, ho, hn, hun, hvn, split, hscheme, vscheme, uuEx,
                     vvEx):
        ps = [self._f3(t, nm) for t, nm in ((uu, "uu"), (vv, "vv"), (ww, "ww"), (ho, "ho"),
                                            (hn, "hn"), (hun, "hun"), (hvn, "hvn"))]
        check(self.lib.gt_uv_advect_3d(self.h, dt, *ps, split, hscheme, vscheme,
                                       self._f3(uuEx, "uuEx"), self._f3(vvEx, "vvEx")))

    # ---- vertical diffusion (temperature.F90:463-551, salinity.F90:346-393) ----
    def salinity_diffusion(self, dt, cnpar, hn, nuh, S):
        check(self.lib.gt_salinity_diffusion(self.h, dt, cnpar, self._f3(hn, "hn"),
                                             self._f3(nuh, "nuh"), self._f3(S, "S")))

    def temperature_diffusion(self, dt, cnpar, hn, nuh, swr, shf, A, g1, g2, dry_z, S, rad, T,
                              swr_bot_refl_frac=-1.0, swr_min_bot_frac=0.01):
        f32 = lambda v: float(np.float32(v))
        p2 = [self._f2(a, nm) for a, nm in ((swr, "swr"), (shf, "shf"), (A, "A"), (g1, "g1"),
                                            (g2, "g2"), (dry_z, "dry_z"))]
        check(self.lib.gt_temperature_diffusion(
            self.h, dt, cnpar, self._f3(hn, "hn"), self._f3(nuh, "nuh"), *p2,
            self._f3(S, "S"), swr_bot_refl_frac, f32(swr_min_bot_frac), self._f3(rad, "rad"),
            self._f3(T, "T")))

    def tridiagonal_batch(self, fi, lt, au, bu, cu, du, value):
        """systems along the last axis: arrays of shape (rows, n) with at
        least lt+1 rows; fi, lt outside 0 <= fi < lt go to the library, which
        rejects them before touching memory"""
        n = au.shape[-1] if au is not None and au.dim() == 2 else 0
        rows = au.shape[0] if n else 0
        if n and 0 <= fi < lt and lt >= rows:
            raise ValueError(f"tridiagonal_batch: lt = {lt} needs at least {lt + 1} rows "
                             f"(got {rows})")
        ps = [_dev(t, nm, shape=(rows, n)) for t, nm in ((au, "au"), (bu, "bu"), (cu, "cu"),
                                                         (du, "du"), (value, "value"))]
        check(self.lib.gt_tridiagonal_batch(self.h, n, fi, lt, *ps))

    def rho_from_theta_unesco80(self, T, S, rho, nonnegsalt=False):
        n = T.numel() if T is not None else 0
        check(self.lib.gt_rho_from_theta_unesco80(self.h, n, _dev(T, "T", numel=n),
                                                  _dev(S, "S", numel=n),
                                                  _dev(rho, "rho", numel=n), int(nonnegsalt)))

    def rho_from_theta(self, s, th, p, dens0, densp=None):
        n = s.numel() if s is not None else 0
        check(self.lib.gt_rho_from_theta(self.h, n, _dev(s, "s", numel=n),
                                         _dev(th, "th", numel=n),
                                         _dev(p, "p", numel=n, optional=True),
                                         _dev(dens0, "dens0", numel=n),
                                         _dev(densp, "densp", numel=n, optional=True)))

    def eosall_from_theta(self, s, th, p, rho_s, rho_th):
        n = s.numel() if s is not None else 0
        check(self.lib.gt_eosall_from_theta(self.h, n, _dev(s, "s", numel=n),
                                            _dev(th, "th", numel=n),
                                            _dev(p, "p", numel=n, optional=True),
                                            _dev(rho_s, "rho_s", numel=n),
                                            _dev(rho_th, "rho_th", numel=n)))

    def ip_blumberg_mellor(self, H, hn, hun, hvn, buoy, zz, idpdx, idpdy):
        ps = [self._f3(t, nm) for t, nm in ((hn, "hn"), (hun, "hun"), (hvn, "hvn"),
                                            (buoy, "buoy"), (zz, "zz"), (idpdx, "idpdx"),
                                            (idpdy, "idpdy"))]
        check(self.lib.gt_ip_blumberg_mellor(self.h, self._f2(H, "H"), *ps))

    def ip_shchepetkin_mcwilliams(self, H, hn, hun, hvn, buoy, sseo, zz, idpdx, idpdy):
        ps = [self._f3(t, nm) for t, nm in ((hn, "hn"), (hun, "hun"), (hvn, "hvn"),
                                            (buoy, "buoy"))]
        check(self.lib.gt_ip_shchepetkin_mcwilliams(
            self.h, self._f2(H, "H"), *ps, self._f2(sseo, "sseo"), self._f3(zz, "zz"),
            self._f3(idpdx, "idpdx"), self._f3(idpdy, "idpdy")))

    def eos_internal_pressure(self, methods, T, S, T0, S0, dtr0, dsr0, H, hn, hun, hvn,
                              sseo=None, zz_bm=None, idpdx_bm=None, idpdy_bm=None, zz_sm=None,
                              idpdx_sm=None, idpdy_sm=None):
        """the linear EOS fused into Blumberg-Mellor (methods & 1) and/or
        Shchepetkin-McWilliams (methods & 2): the gradients of eos_linear
        followed by ip_blumberg_mellor / ip_shchepetkin_mcwilliams, bit for
        bit, without storing the buoyancy"""
        bm, sm = bool(methods & 1), bool(methods & 2)
        ps = [self._f3(t, nm) for t, nm in ((T, "T"), (S, "S"))]
        hs = [self._f3(t, nm) for t, nm in ((hn, "hn"), (hun, "hun"), (hvn, "hvn"))]
        outs = [self._f3(t, nm, optional=not on) for t, nm, on in (
            (zz_bm, "zz_bm", bm), (idpdx_bm, "idpdx_bm", bm), (idpdy_bm, "idpdy_bm", bm),
            (zz_sm, "zz_sm", sm), (idpdx_sm, "idpdx_sm", sm), (idpdy_sm, "idpdy_sm", sm))]
        check(self.lib.gt_eos_internal_pressure(
            self.h, methods, *ps, T0, S0, dtr0, dsr0, self._f2(H, "H"), *hs,
            self._f2(sseo, "sseo", optional=not sm), *outs))

    def eqstate_internal_pressure(self, eqstate_method, methods, T, S, rho, H, hn, hun, hvn,
                                  sseo=None, zz_bm=None, idpdx_bm=None, idpdy_bm=None,
                                  zz_sm=None, idpdx_sm=None, idpdy_sm=None, nonnegsalt=False,
                                  T0=10.0, S0=33.75, dtr0=-0.17, dsr0=0.78):
        """do_eqstate of any method (eqstate.F90:94-261) fused into
        Blumberg-Mellor and/or Shchepetkin-McWilliams: the gradients of
        do_eqstate followed by ip_blumberg_mellor / ip_shchepetkin_mcwilliams,
        bit for bit; rho is read on land columns (methods 2 and 3) and never
        written.  The linear constants default to the reference's namelist
        values, FP32 literals (rounded here)."""
        f32 = lambda v: float(np.float32(v))
        bm, sm = bool(methods & 1), bool(methods & 2)
        ps = [self._f3(t, nm) for t, nm in ((T, "T"), (S, "S"))]
        hs = [self._f3(t, nm) for t, nm in ((hn, "hn"), (hun, "hun"), (hvn, "hvn"))]
        outs = [self._f3(t, nm, optional=not on) for t, nm, on in (
            (zz_bm, "zz_bm", bm), (idpdx_bm, "idpdx_bm", bm), (idpdy_bm, "idpdy_bm", bm),
            (zz_sm, "zz_sm", sm), (idpdx_sm, "idpdx_sm", sm), (idpdy_sm, "idpdy_sm", sm))]
        check(self.lib.gt_eqstate_internal_pressure(
            self.h, eqstate_method, int(nonnegsalt), methods, *ps,
            self._f3(rho, "rho", optional=eqstate_method == 1), f32(T0), f32(S0), f32(dtr0),
            f32(dsr0), self._f2(H, "H"), *hs, self._f2(sseo, "sseo", optional=not sm), *outs))

    def do_internal_pressure(self, ip_method, H, hn, hun, hvn, buoy, sseo, idpdx, idpdy):
        ps = [self._f3(t, nm) for t, nm in ((hn, "hn"), (hun, "hun"), (hvn, "hvn"),
                                            (buoy, "buoy"))]
        check(self.lib.gt_do_internal_pressure(
            self.h, ip_method, self._f2(H, "H"), *ps,
            self._f2(sseo, "sseo", optional=ip_method != SHCHEPETKIN_MCWILLIAMS),
            self._f3(idpdx, "idpdx"), self._f3(idpdy, "idpdy")))

    def set_substr_ini_press(self, on=True):
        """gt_set_substr_ini_press: the SUBSTR_INI_PRESS build of
        do_internal_pressure (internal_pressure.F90:235-258)"""
        check(self.lib.gt_set_substr_ini_press(self.h, int(on)))

    # ---- ghost cells -----------------------------------------------------------
    def update_3d_halo(self, f1, f2, tag=H_TAG, mirror=-1):
        check(self.lib.gt_update_3d_halo(self.h, self._f3(f1, "f1"), self._f3(f2, "f2"), tag,
                                         mirror))

    def update_2d_halo(self, f1, f2, tag=H_TAG, mirror=-1):
        check(self.lib.gt_update_2d_halo(self.h, self._f2(f1, "f1"), self._f2(f2, "f2"), tag,
                                         mirror))

    def wait_halo(self, tag=H_TAG):
        check(self.lib.gt_wait_halo(self.h, tag))

    # ---- budgets ---------------------------------------------------------------
    def budget(self, h, f=None):
        """(mass, volume); f may be an FP64 or FP32 field"""
        import torch
        m, v = C.c_double(), C.c_double()
        f32 = f is not None and f.dtype == torch.float32
        fn = self.lib.gt_budget_f32 if f32 else self.lib.gt_budget
        check(fn(self.h, self._f3(h, "h"),
                 self._f3(f, "f", True, dtype=torch.float32 if f32 else None),
                 C.byref(m), C.byref(v)))
        return m.value, v.value

    def budget_batch(self, h, f, ntracers=None, out=None):
        """the FP64 budgets of a set of FP32 tracers (ntracers, kmax+1, NY, NX)
        sharing h, in one pass: a device tensor of ntracers + 1 doubles (the
        masses, then the volume), asynchronous on the context's stream"""
        import torch
        nt = f.shape[0] if ntracers is None else ntracers
        if out is None:
            out = torch.empty(nt + 1, dtype=torch.float64, device=h.device)
        check(self.lib.gt_budget_batch_f32(self.h, nt, self._f3(h, "h"),
                                           self._set(f, "f", nt, torch.float32),
                                           _dev(out, "out", numel=nt + 1)))
        return out

    # ---- reverse mode -------------------------------------------------------------
    def _vel(self, uu, vv, ww, hu, hv, ho):
        return [self._f3(t, nm) for t, nm in ((uu, "uu"), (vv, "vv"), (ww, "ww"), (hu, "hu"),
                                              (hv, "hv"), (ho, "ho"))]

    def adv3d_adjoint(self, nsteps, ckpt_every, dt, f0, lam_final, lam0, uu, vv, ww, hu, hv, ho,
                      hscheme, vscheme, AH=-1.0, tag=H_TAG, f_final=None):
        check(self.lib.gt_adv3d_adjoint(self.h, nsteps, ckpt_every, dt, self._f3(f0, "f0"),
                                        self._f3(f_final, "f_final", True),
                                        self._f3(lam_final, "lam_final"),
                                        self._f3(lam0, "lam0"),
                                        *self._vel(uu, vv, ww, hu, hv, ho), hscheme,
                                        vscheme, AH, tag))

    def adv3d_adjoint_f32(self, ntracers, nsteps, ckpt_every, dt, f0, lam_final, lam0, uu, vv,
                          ww, hu, hv, ho, hscheme, vscheme, AH=-1.0, tag=H_TAG, f_final=None):
        """gt_adv3d_adjoint_f32: f0, lam_final, lam0 (and f_final) are FP32
        tensors of shape (ntracers, kmax+1, NY, NX)"""
        import torch
        st = lambda t, nm, o=False: self._set(t, nm, ntracers, torch.float32, o)
        check(self.lib.gt_adv3d_adjoint_f32(self.h, ntracers, nsteps, ckpt_every, dt,
                                            st(f0, "f0"), st(f_final, "f_final", True),
                                            st(lam_final, "lam_final"), st(lam0, "lam0"),
                                            *self._vel(uu, vv, ww, hu, hv, ho),
                                            hscheme, vscheme, AH, tag))

    def adv3d_adjoint_seq(self, ntracers, ckpt_every, dt, f0, lam_final, lam0, steps, hscheme,
                          vscheme, AH=-1.0, tag=H_TAG, f_final=None):
        """gt_adv3d_adjoint_seq (FP64 tensors, ntracers = 1) or _seq_f32 (FP32,
        shape (ntracers, kmax+1, NY, NX)); steps: one (uu, vv, ww, hu, hv, ho)
        tuple of device tensors per step"""
        import torch
        n = len(steps)
        for t in range(n):
            if len(steps[t]) != 6:
                raise ValueError("adv3d_adjoint_seq: each step needs (uu, vv, ww, hu, hv, ho)")
            self._vel(*steps[t])
        arrs = []
        for q in range(6):
            a = (C.c_void_p * n)(*[steps[t][q].data_ptr() for t in range(n)])
            arrs.append(a)
        ptrs = [C.cast(a, C.POINTER(C.c_void_p)) for a in arrs]
        self._seq_keep = arrs  # alive until the call returns
        f32 = f0 is not None and f0.dtype == torch.float32
        dt_ = torch.float32 if f32 else torch.float64
        st = lambda t, nm, o=False: self._set(t, nm, ntracers, dt_, o)
        if f32:
            check(self.lib.gt_adv3d_adjoint_seq_f32(self.h, ntracers, n, ckpt_every, dt,
                                                    st(f0, "f0"), st(f_final, "f_final", True),
                                                    st(lam_final, "lam_final"), st(lam0, "lam0"),
                                                    *ptrs, hscheme, vscheme, AH, tag))
        else:
            if ntracers != 1:
                raise ValueError("adv3d_adjoint_seq: FP64 takes one tracer")
            check(self.lib.gt_adv3d_adjoint_seq(self.h, n, ckpt_every, dt, st(f0, "f0"),
                                                st(f_final, "f_final", True),
                                                st(lam_final, "lam_final"), st(lam0, "lam0"),
                                                *ptrs, hscheme, vscheme, AH, tag))

    def ip_blumberg_mellor_adjoint(self, hun, hvn, zz, idpdx_bar, idpdy_bar, buoy_bar):
        ps = [self._f3(t, nm) for t, nm in ((hun, "hun"), (hvn, "hvn"), (zz, "zz"),
                                            (idpdx_bar, "idpdx_bar"), (idpdy_bar, "idpdy_bar"),
                                            (buoy_bar, "buoy_bar"))]
        check(self.lib.gt_ip_blumberg_mellor_adjoint(self.h, *ps))

    def ip_shchepetkin_mcwilliams_adjoint(self, hn, hun, hvn, buoy, sseo, zz, idpdx_bar,
                                          idpdy_bar, buoy_bar):
        ps = [self._f3(t, nm) for t, nm in ((hn, "hn"), (hun, "hun"), (hvn, "hvn"),
                                            (buoy, "buoy"))]
        qs = [self._f3(t, nm) for t, nm in ((zz, "zz"), (idpdx_bar, "idpdx_bar"),
                                            (idpdy_bar, "idpdy_bar"), (buoy_bar, "buoy_bar"))]
        check(self.lib.gt_ip_shchepetkin_mcwilliams_adjoint(self.h, *ps, self._f2(sseo, "sseo"),
                                                            *qs))

    def eos_linear_adjoint(self, buoy_bar, T_bar, S_bar, dtr0, dsr0):
        check(self.lib.gt_eos_linear_adjoint(self.h, self._f3(buoy_bar, "buoy_bar"),
                                             self._f3(T_bar, "T_bar", True),
                                             self._f3(S_bar, "S_bar", True), dtr0, dsr0))

    def set_tuning(self, fused=1, tile=0, group=3):
        check(self.lib.gt_set_tuning(self.h, int(fused), int(tile), int(group)))

    def set_adjoint_mode(self, keep_states=-1):
        """gt_set_adjoint_mode: 1 keeps each segment's sub-step states, 0
        recomputes them per reverse step, -1 decides by free memory"""
        check(self.lib.gt_set_adjoint_mode(self.h, int(keep_states)))


def exchange_tiles(px, py, ni, nj, nk, tiles, stream=None):
    """gt_exchange_tiles on a list of device tensors: tile t = ty*px + tx has
    shape (>= nk, nj[t]+4, ni[t]+4)"""
    lib = load_library()
    n = px * py
    if len(tiles) != n or len(ni) != n or len(nj) != n:
        raise ValueError("exchange_tiles: need px*py tiles and sizes")
    for q, t in enumerate(tiles):
        _dev(t, f"tile {q}")
        if t.ndim != 3 or t.shape[1:] != (nj[q] + 4, ni[q] + 4) or t.shape[0] < nk:
            raise ValueError(f"exchange_tiles: tile {q} has shape {tuple(t.shape)}, needs "
                             f"(>= {nk}, {nj[q] + 4}, {ni[q] + 4})")
    ni_a = (C.c_int * n)(*ni)
    nj_a = (C.c_int * n)(*nj)
    tp = (C.c_void_p * n)(*[t.data_ptr() for t in tiles])
    handle = getattr(stream, "cuda_stream", stream)
    check(lib.gt_exchange_tiles(px, py, ni_a, nj_a, nk, tp,
                                C.c_void_p(handle) if handle else None))


def _batch_array(domains, ntracers):
    """gt_batch_domain[] from dicts, each tensor checked against its
    context: one tracer dtype for the whole batch, f_in/f_out of shape
    (ntracers, kmax+1, NY, NX), FP64 velocities and heights, one context per
    domain; optional hires and budget (ntracers + 1 FP64 values: the
    masses of the result, then the volume)"""
    import torch
    if not domains:
        raise ValueError("adv3d_batch: no domains")
    dtype = domains[0]["f_in"].dtype
    if dtype not in (torch.float32, torch.float64):
        raise ValueError(f"adv3d_batch: tracers must be FP32 or FP64 (got {dtype})")
    seen = set()
    n = len(domains)
    arr = (BatchDomain * n)()
    for q, d in enumerate(domains):
        ctx = d["ctx"]
        if id(ctx) in seen:
            raise ValueError(f"adv3d_batch: domain {q} repeats a context (one per domain)")
        seen.add(id(ctx))
        shp = (ntracers,) + ctx.shape3
        for key in ("f_in", "f_out"):
            _dev(d[key], f"domain {q} {key}", dtype=dtype, numel=int(np.prod(shp)))
        if d["f_in"].data_ptr() == d["f_out"].data_ptr():
            raise ValueError(f"adv3d_batch: domain {q} f_in and f_out must differ")
        vel = [ctx._f3(d[k], f"domain {q} {k}") for k in ("uu", "vv", "ww", "hu", "hv", "ho")]
        hires = ctx._f3(d.get("hires"), f"domain {q} hires", True)
        budget = _dev(d.get("budget"), f"domain {q} budget", numel=ntracers + 1,
                      optional=True)
        arr[q] = BatchDomain(ctx.h.value, d["f_in"].data_ptr(), d["f_out"].data_ptr(),
                             *(v.value for v in vel), hires.value if hires is not None else None,
                             budget.value if budget is not None else None)
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
    arr = _batch_array(domains, ntracers)
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
    ho and optional hires and budget (ntracers + 1 FP64 values, formed in
    the step).  The dtype of f_in picks the FP32 or FP64 kernel."""
    import torch
    lib = load_library()
    n = len(domains)
    arr = _batch_array(domains, ntracers)
    f32 = domains[0]["f_in"].dtype == torch.float32
    fn = lib.gt_adv3d_batch_f32 if f32 else lib.gt_adv3d_batch_f64
    handle = getattr(stream, "cuda_stream", stream)
    check(fn(n, arr, ntracers, dt, hscheme, vscheme, AH, tag,
             C.c_void_p(handle) if handle else None))
