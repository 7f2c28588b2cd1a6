"""ctypes binding of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to check results. The product never imports it.
"""
import ctypes as C
import os

import numpy as np

from kapsel3_amd._abi import CTime, KhParams, PARTICLE_DTYPE, FIELD, FN, ptr

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.kr_create.restype = vp
        L.kr_create.argtypes = [C.POINTER(KhParams), C.c_int]
        L.kr_destroy.argtypes = [vp]
        for name in ("kr_set_particles", "kr_get_particles"):
            getattr(L, name).argtypes = [vp, vp, C.c_int]
        for name in ("kr_set_fluid", "kr_get_fluid"):
            getattr(L, name).argtypes = [vp, vp, vp, vp]
        L.kr_set_time.argtypes = [vp, C.POINTER(CTime)]
        L.kr_get_time.argtypes = [vp, C.POINTER(CTime)]
        L.kr_get_field.argtypes = [vp, C.c_int, vp]
        L.kr_set_field.argtypes = [vp, C.c_int, vp]
        L.kr_init_fluid.argtypes = [vp]
        L.kr_step.argtypes = [vp, C.c_int]
        L.kr_apply.argtypes = [vp, C.c_int]
        L.kr_A2a_k.argtypes = [vp, vp]
        L.kr_A_k2a.argtypes = [vp, vp]
        L.kr_fft_forward.argtypes = [C.c_int, C.c_int, C.c_int, vp]
        L.kr_fft_inverse.argtypes = [C.c_int, C.c_int, C.c_int, vp]
        L.kr_phi.restype = C.c_double
        L.kr_phi.argtypes = [C.c_double] * 4
        L.kr_squirmer_coord.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.kr_squirmer_coord.restype = None
        L.kr_sekibun_count.argtypes = [vp]
        L.kr_sekibun_cell.argtypes = [vp, vp]
        L.kr_philox4x32_10.argtypes = [vp, vp, vp]
        L.kr_noise_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, vp]
        L.kr_set_threads.argtypes = [C.c_int]
        L.kr_set_rigid.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.kr_get_rigid.argtypes = [vp, vp]
        L.kr_get_rigid.restype = None
        L.kr_set_mutation.argtypes = [C.c_int]
        L.kr_set_mutation.restype = None
        L.kr_set_shear.argtypes = [vp, C.c_double, C.c_int, C.c_double]
        L.kr_get_shear_state.argtypes = [vp, vp, vp]
        L.kr_get_shear_state.restype = None
        L.kr_set_shear_state.argtypes = [vp, C.c_double, C.c_double]
        L.kr_set_shear_state.restype = None
        L.kr_le_apply.argtypes = [vp, C.c_int]
        L.kr_get_shear_record.argtypes = [vp, vp]
        L.kr_get_shear_record.restype = None
        L.kr_get_shear_diag.argtypes = [vp, vp]
        L.kr_get_shear_diag.restype = None
        _LIB = L
    return _LIB


# single functions of the sheared step (enum in kref.h)
LE_FN = {name: i for i, name in enumerate([
    "oblique2u", "u2oblique", "reset_up", "reset_down", "zeta_k2u", "u2zeta_k", "solenoidal_u",
    "ns_solver", "shear_rate_eff", "position_euler", "make_phi", "make_up", "f_hydro", "lj", "reynolds"])}
# fields of the per-step stress record (Mean_shear_stress, sp_3d_ns.h:247-330)
LE_RECORD = ["time", "shear_rate", "degree_oblique", "strain", "lj_dev_stress", "shear_stress_old",
             "shear_stress_new", "reynolds_stress", "fluid_stress", "interfacial_stress", "apparent_stress",
             "viscosity"]


# planted mistakes (kr_set_mutation, kref.h), in enum order
MUTATIONS = ["none", "fh_pre_ns_u", "omega_r_sign", "fh_no_max", "up_no_max", "no_fix_cell", "adv_sign",
             "ab2_fh_weight", "no_euler_start", "fh_old_x", "fh_sign", "ukdc_mean", "slip_relax", "slip_b2_full",
             "slip_force_sign", "slip_final_ustar", "slip_rotlet", "slip_no_reset"]


class mutation:
    """Context manager: the oracle commits the named planted mistake inside
    the block (tests of the pins only)."""

    def __init__(self, name):
        self.m = MUTATIONS.index(name)

    def __enter__(self):
        lib().kr_set_mutation(self.m)
        return self

    def __exit__(self, *exc):
        lib().kr_set_mutation(0)


def fft_forward(a):
    """A2a_k on a reference-layout array of shape (nx, ny, nz_)."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    nx, ny, nz_ = a.shape
    nz = nz_ - 2
    if lib().kr_fft_forward(nx, ny, nz, a.ctypes.data) != 0:
        raise ValueError("unsupported size")
    return a


def fft_inverse(a):
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    nx, ny, nz_ = a.shape
    if lib().kr_fft_inverse(nx, ny, nz_ - 2, a.ctypes.data) != 0:
        raise ValueError("unsupported size")
    return a


def phi(r, radius, dx, xi):
    return lib().kr_phi(r, radius, dx, xi)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().kr_philox4x32_10(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return o


def noise_normals(seed, stream, particle, ts):
    o = np.zeros(6)
    lib().kr_noise_normals(seed, stream, particle, ts, o.ctypes.data)
    return o


def set_threads(n):
    lib().kr_set_threads(int(n))


class Run:
    """One oracle simulation; mirrors kapsel3_amd.Run."""

    def __init__(self, params, nparticles):
        self.params = params
        self.np = nparticles
        self.h = lib().kr_create(C.byref(params.c), nparticles)
        if not self.h:
            raise ValueError("kr_create rejected the parameters")
        self.shape = (params.c.nx, params.c.ny, params.nz_)

    def close(self):
        if self.h:
            lib().kr_destroy(self.h)
            self.h = None

    __del__ = close

    def set_particles(self, p):
        p = np.ascontiguousarray(p, dtype=PARTICLE_DTYPE)
        if lib().kr_set_particles(self.h, p.ctypes.data, len(p)) != 0:
            raise ValueError("bad particles")

    def get_particles(self):
        p = np.zeros(self.np, dtype=PARTICLE_DTYPE)
        lib().kr_get_particles(self.h, p.ctypes.data, self.np)
        return p

    def set_fluid(self, z0, z1, ukdc):
        z0 = np.ascontiguousarray(z0, dtype=np.float64)
        z1 = np.ascontiguousarray(z1, dtype=np.float64)
        uk = np.ascontiguousarray(ukdc, dtype=np.float64)
        lib().kr_set_fluid(self.h, z0.ctypes.data, z1.ctypes.data, uk.ctypes.data)

    def get_fluid(self):
        z0 = np.zeros(self.shape)
        z1 = np.zeros(self.shape)
        uk = np.zeros(3)
        lib().kr_get_fluid(self.h, z0.ctypes.data, z1.ctypes.data, uk.ctypes.data)
        return z0, z1, uk

    def get_time(self):
        t = CTime()
        lib().kr_get_time(self.h, C.byref(t))
        return t

    def set_time(self, t):
        lib().kr_set_time(self.h, C.byref(t))

    def get_field(self, name):
        a = np.zeros(self.shape)
        lib().kr_get_field(self.h, FIELD[name], a.ctypes.data)
        return a

    def set_field(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        lib().kr_set_field(self.h, FIELD[name], a.ctypes.data)

    def init_fluid(self):
        lib().kr_init_fluid(self.h)

    def step(self, n=1):
        lib().kr_step(self.h, n)

    def apply(self, fn):
        if lib().kr_apply(self.h, FN[fn]) != 0:
            raise ValueError(fn)

    def set_shear(self, rate, ac=False, frequency=0.0):
        """Switch to the Lees-Edwards step at this shear rate."""
        if lib().kr_set_shear(self.h, float(rate), int(bool(ac)), float(frequency)) != 0:
            raise ValueError("kr_set_shear failed")

    def shear_state(self):
        g, s = C.c_double(), C.c_double()
        lib().kr_get_shear_state(self.h, C.byref(g), C.byref(s))
        return g.value, s.value

    def set_shear_state(self, degree_oblique, shear_rate_eff):
        lib().kr_set_shear_state(self.h, float(degree_oblique), float(shear_rate_eff))

    def shear_record(self):
        o = np.zeros(len(LE_RECORD))
        lib().kr_get_shear_record(self.h, o.ctypes.data)
        return dict(zip(LE_RECORD, o))

    def shear_diag(self):
        o = np.zeros(4)
        lib().kr_get_shear_diag(self.h, o.ctypes.data)
        return o

    def le_apply(self, fn):
        if lib().kr_le_apply(self.h, LE_FN[fn]) != 0:
            raise ValueError(fn)

    def A2a_k(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        lib().kr_A2a_k(self.h, a.ctypes.data)
        return a

    def A_k2a(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).copy()
        lib().kr_A_k2a(self.h, a.ctypes.data)
        return a

    def set_rigid(self, cum, free=None, v0=None, w0=None):
        """Rigid bodies: body k = particles cum[k]..cum[k+1]-1 (kr_set_rigid)."""
        cum = np.ascontiguousarray(cum, dtype=np.int32)
        fr = None if free is None else np.ascontiguousarray(free, dtype=np.int32)
        v = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        w = None if w0 is None else np.ascontiguousarray(w0, dtype=np.float64)
        self._rigid = (cum, fr, v, w)
        if lib().kr_set_rigid(self.h, len(cum) - 1, cum.ctypes.data, None if fr is None else fr.ctypes.data,
                              None if v is None else v.ctypes.data, None if w is None else w.ctypes.data) != 0:
            raise ValueError("kr_set_rigid rejected the bodies")
        self.nb = len(cum) - 1

    def get_rigid(self):
        from kapsel3_amd._abi import RIGID_DTYPE
        out = np.zeros(self.nb, dtype=RIGID_DTYPE)
        lib().kr_get_rigid(self.h, out.ctypes.data)
        return out

    def sekibun(self):
        n = lib().kr_sekibun_count(self.h)
        out = np.zeros((n, 3), dtype=np.int32)
        lib().kr_sekibun_cell(self.h, out.ctypes.data)
        return out
