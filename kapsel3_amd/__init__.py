"""MI355X-native smoothed-profile-method (SPM) step of KAPSEL.

Python binding of the C ABI in include/kapsel_hip.h (libkapsel_hip.so, built
in-tree by kapsel3_amd.build). The simulation state lives in HBM inside a
`Run`; this module only moves parameters, particles and fields across the
boundary and enqueues steps. There is no CPU fallback: without the HIP
library every call raises.
"""
import ctypes as C
import os

import numpy as np

from ._abi import (CTime, FIELD, FN, FP64, JAX_NONE, JAX_X, JAX_Y, JAX_Z, JP_MOTOR, JP_NONE,  # noqa: F401
                   JP_OBSTACLE, JP_SLIP, MIXED, PARTICLE_DTYPE, KhParams, KhSpecies, Params, make_species,
                   new_particles)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkapsel_hip.so")
_LIB = None

EXPORTS = [
    "kh_run_create", "kh_run_destroy", "kh_set_particles", "kh_get_particles", "kh_set_fluid", "kh_get_fluid",
    "kh_set_time", "kh_get_time", "kh_get_field", "kh_set_field", "kh_init_fluid", "kh_step", "kh_apply",
    "kh_sync", "kh_A2a_k", "kh_A_k2a", "kh_set_step_path", "kh_set_graph", "kh_timing_enable", "kh_timing_reset", "kh_timing_get",
    "kh_timing_count", "kh_timing_name", "kh_strerror", "kh_last_error",
]


class KhError(RuntimeError):
    pass


def lib():
    """Loads libkapsel_hip.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise KhError(f"{LIB_PATH} is not built; run `python -m kapsel3_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.kh_run_create.argtypes = [C.POINTER(KhParams), C.c_int, C.POINTER(vp)]
        L.kh_run_destroy.argtypes = [vp]
        for n in ("kh_set_particles", "kh_get_particles"):
            getattr(L, n).argtypes = [vp, vp, C.c_int]
        for n in ("kh_set_fluid", "kh_get_fluid"):
            getattr(L, n).argtypes = [vp, vp, vp, vp]
        L.kh_set_time.argtypes = [vp, C.POINTER(CTime)]
        L.kh_get_time.argtypes = [vp, C.POINTER(CTime)]
        L.kh_get_field.argtypes = [vp, C.c_int, vp]
        L.kh_set_field.argtypes = [vp, C.c_int, vp]
        L.kh_init_fluid.argtypes = [vp]
        L.kh_step.argtypes = [vp, C.c_int]
        L.kh_apply.argtypes = [vp, C.c_int]
        L.kh_sync.argtypes = [vp]
        L.kh_set_step_path.argtypes = [vp, C.c_int]
        L.kh_set_graph.argtypes = [vp, C.c_int]
        L.kh_A2a_k.argtypes = [vp, vp]
        L.kh_A_k2a.argtypes = [vp, vp]
        L.kh_timing_enable.argtypes = [vp, C.c_int]
        L.kh_timing_reset.argtypes = [vp]
        L.kh_timing_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.kh_timing_count.argtypes = [vp]
        L.kh_timing_name.argtypes = [vp, C.c_int]
        L.kh_timing_name.restype = C.c_char_p
        L.kh_strerror.restype = C.c_char_p
        L.kh_strerror.argtypes = [C.c_int]
        L.kh_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(status):
    if status < 0:
        L = lib()
        raise KhError(f"{L.kh_strerror(status).decode()}: {L.kh_last_error().decode()}")
    return status


class Run:
    """One device-resident SPM simulation (kh_run)."""

    def __init__(self, params, nparticles):
        self.params = params
        self.np = int(nparticles)
        h = C.c_void_p()
        _check(lib().kh_run_create(C.byref(params.c), self.np, C.byref(h)))
        self.h = h
        self.shape = (params.c.nx, params.c.ny, params.nz_)

    def close(self):
        if getattr(self, "h", None):
            lib().kh_run_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state transfer
    def set_particles(self, p):
        p = np.ascontiguousarray(p, dtype=PARTICLE_DTYPE)
        _check(lib().kh_set_particles(self.h, p.ctypes.data, len(p)))

    def get_particles(self, into=None):
        p = new_particles(self.np) if into is None else into
        _check(lib().kh_get_particles(self.h, p.ctypes.data, self.np))
        return p

    def set_fluid(self, z0, z1, ukdc):
        z0 = np.ascontiguousarray(z0, dtype=np.float64)
        z1 = np.ascontiguousarray(z1, dtype=np.float64)
        uk = np.ascontiguousarray(ukdc, dtype=np.float64)
        _check(lib().kh_set_fluid(self.h, z0.ctypes.data, z1.ctypes.data, uk.ctypes.data))

    def get_fluid(self):
        z0, z1, uk = np.zeros(self.shape), np.zeros(self.shape), np.zeros(3)
        _check(lib().kh_get_fluid(self.h, z0.ctypes.data, z1.ctypes.data, uk.ctypes.data))
        return z0, z1, uk

    def get_time(self):
        t = CTime()
        _check(lib().kh_get_time(self.h, C.byref(t)))
        return t

    def set_time(self, t):
        _check(lib().kh_set_time(self.h, C.byref(t)))

    def get_field(self, name):
        a = np.zeros(self.shape)
        _check(lib().kh_get_field(self.h, FIELD[name], a.ctypes.data))
        return a

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _check(lib().kh_set_field(self.h, FIELD[name], a.ctypes.data))

    # ---- the step
    def init_fluid(self):
        _check(lib().kh_init_fluid(self.h))

    def step(self, n=1):
        _check(lib().kh_step(self.h, int(n)))

    def apply(self, fn):
        _check(lib().kh_apply(self.h, FN[fn]))

    def set_step_path(self, path):
        """'fused' (default) or 'reference' (the reference's operation order)."""
        _check(lib().kh_set_step_path(self.h, {"fused": 0, "reference": 1}[path]))

    def set_graph(self, on=True):
        """Run steady-state steps as one captured HIP graph (default on)."""
        _check(lib().kh_set_graph(self.h, 1 if on else 0))

    def sync(self):
        _check(lib().kh_sync(self.h))

    def A2a_k(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        _check(lib().kh_A2a_k(self.h, a.ctypes.data))
        return a

    def A_k2a(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        _check(lib().kh_A_k2a(self.h, a.ctypes.data))
        return a

    # ---- measurement
    def timing(self, on=True):
        _check(lib().kh_timing_enable(self.h, 1 if on else 0))

    def timing_reset(self):
        _check(lib().kh_timing_reset(self.h))

    def timing_stats(self):
        L = lib()
        n = _check(L.kh_timing_count(self.h))
        out = {}
        for i in range(n):
            name = L.kh_timing_name(self.h, i)
            cnt, ms = C.c_int(), C.c_double()
            _check(L.kh_timing_get(self.h, name, C.byref(cnt), C.byref(ms)))
            out[name.decode()] = (cnt.value, ms.value)
        return out
