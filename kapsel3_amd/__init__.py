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
                   JP_OBSTACLE, JP_SLIP, MIXED, PARTICLE_DTYPE, RIGID_DTYPE, KhParams, KhSpecies, Params,
                   make_species, new_particles)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkapsel_hip.so")
_LIB = None

EXPORTS = [
    "kh_run_create", "kh_run_destroy", "kh_set_particles", "kh_get_particles", "kh_set_fluid", "kh_get_fluid",
    "kh_set_time", "kh_get_time", "kh_get_field", "kh_set_field", "kh_init_fluid", "kh_step", "kh_apply",
    "kh_sync", "kh_A2a_k", "kh_A_k2a", "kh_set_step_path", "kh_set_graph", "kh_timing_enable", "kh_timing_reset", "kh_timing_get",
    "kh_timing_count", "kh_timing_name", "kh_strerror", "kh_last_error", "kh_set_trajectory", "kh_get_trajectory",
    "kh_set_output", "kh_get_output", "kh_run_bytes", "kh_query", "kh_server_create", "kh_server_submit", "kh_server_poll", "kh_server_wait",
    "kh_server_run_times", "kh_server_get_stats", "kh_server_destroy", "kh_set_shear", "kh_get_shear_state",
    "kh_set_shear_state", "kh_get_shear_record", "kh_set_shear_history", "kh_get_shear_history",
    "kh_server_submit_output", "kh_set_rigid", "kh_get_rigid", "kh_A2a_k_noalloc",
    "kh_get_slip_iterations",
]

# fields of the sheared step's stress record (Mean_shear_stress, sp_3d_ns.h:247-330)
SHEAR_RECORD = ["time", "shear_rate", "degree_oblique", "strain", "lj_dev_stress", "shear_stress_old",
                "shear_stress_new", "reynolds_stress", "fluid_stress", "interfacial_stress", "apparent_stress",
                "viscosity"]


class ServerConfig(C.Structure):
    _fields_ = [("max_active", C.c_int), ("chunk_steps", C.c_int), ("hbm_fraction", C.c_double),
                ("hbm_budget", C.c_longlong), ("batch_slots", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_longlong), ("completed", C.c_longlong), ("failed", C.c_longlong),
                ("queued", C.c_longlong), ("active", C.c_longlong), ("peak_active", C.c_int),
                ("peak_bytes", C.c_longlong), ("budget", C.c_longlong),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("wall_s", C.c_double), ("runs_per_s", C.c_double), ("steps_per_s", C.c_double),
                ("grid_point_steps_per_s", C.c_double)]


# One particle of an Output_udf frame (output.cxx:268-375), members in the
# record's order; forces and torques are the previous step's.
FRAME_PARTICLE_DTYPE = np.dtype([
    ("R", "f8", 3), ("R_raw", "f8", 3), ("v", "f8", 3), ("q", "f8", 4), ("omega", "f8", 3),
    ("f_hydro", "f8", 3), ("torque_hydro", "f8", 3), ("f_r", "f8", 3), ("torque_r", "f8", 3),
    ("f_slip", "f8", 3), ("torque_slip", "f8", 3)])
assert FRAME_PARTICLE_DTYPE.itemsize == 34 * 8


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
        L.kh_A2a_k_noalloc.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.kh_get_slip_iterations.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.kh_set_rigid.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.kh_get_rigid.argtypes = [vp, vp, C.c_int]
        L.kh_timing_enable.argtypes = [vp, C.c_int]
        L.kh_timing_reset.argtypes = [vp]
        L.kh_timing_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.kh_timing_count.argtypes = [vp]
        L.kh_timing_name.argtypes = [vp, C.c_int]
        L.kh_timing_name.restype = C.c_char_p
        L.kh_strerror.restype = C.c_char_p
        L.kh_strerror.argtypes = [C.c_int]
        L.kh_last_error.restype = C.c_char_p
        L.kh_set_trajectory.argtypes = [vp, C.c_int]
        L.kh_get_trajectory.argtypes = [vp, vp, C.c_int]
        L.kh_set_output.argtypes = [vp, C.c_int, C.c_int]
        L.kh_get_output.argtypes = [vp, vp, C.c_int]
        L.kh_run_bytes.argtypes = [C.POINTER(KhParams), C.c_int]
        L.kh_run_bytes.restype = C.c_longlong
        L.kh_query.argtypes = [vp]
        L.kh_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp)]
        L.kh_server_submit.argtypes = [vp, C.POINTER(KhParams), vp, C.c_int, C.c_int, vp, C.POINTER(C.c_longlong)]
        L.kh_server_submit_output.argtypes = [vp, C.POINTER(KhParams), vp, C.c_int, C.c_int, vp, C.c_int, vp,
                                              C.POINTER(C.c_longlong)]
        L.kh_server_poll.argtypes = [vp, C.c_longlong]
        L.kh_server_wait.argtypes = [vp, C.c_longlong, C.c_double]
        L.kh_server_run_times.argtypes = [vp, C.c_longlong, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.kh_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
        L.kh_server_destroy.argtypes = [vp]
        L.kh_set_shear.argtypes = [vp, C.c_double, C.c_int, C.c_double]
        L.kh_get_shear_state.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.kh_set_shear_state.argtypes = [vp, C.c_double, C.c_double]
        L.kh_get_shear_record.argtypes = [vp, vp]
        L.kh_set_shear_history.argtypes = [vp, C.c_int]
        L.kh_get_shear_history.argtypes = [vp, vp, C.c_int]
        _LIB = L
    return _LIB


def _check(status):
    if status < 0:
        L = lib()
        raise KhError(f"{L.kh_strerror(status).decode()}: {L.kh_last_error().decode()}")
    return status


def A2a_k_noalloc(a):
    """Forward FP64 transform of a host array in the reference layout
    (NX, NY, NZ+2) without a run and without any device allocation
    (kh_A2a_k_noalloc: fixed device scratch of KH_NOALLOC_MAX_BYTES)."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    if a.ndim != 3:
        raise KhError("A2a_k_noalloc takes an (NX, NY, NZ+2) array")
    nx, ny, nz2 = a.shape
    _check(lib().kh_A2a_k_noalloc(a.ctypes.data, nx, ny, nz2 - 2))
    return a


class Run:
    """One device-resident SPM simulation (kh_run)."""

    def __init__(self, params, nparticles):
        self.params = params
        self.np = int(nparticles)
        h = C.c_void_p()
        _check(lib().kh_run_create(C.byref(params.c), self.np, C.byref(h)))
        self.h = h
        self.shape = (params.c.nx, params.c.ny, params.nz_)
        self._traj_steps = 0

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

    # ---- restart
    def save_state(self, path):
        """Write the restart state to an .npz file: the data Save_Restart_udf
        saves (resume.cxx:10-74: zeta, uk_dc, every Particle member including
        the AB2 history, jikan) plus the grid and precision it belongs to.
        Thermal noise is keyed on (seed, stream_id, ts, particle), so a run
        resumed from this state draws the same stream."""
        z0, z1, uk = self.get_fluid()
        t = self.get_time()
        c = self.params.c
        extra = {}
        sh = getattr(self, "_shear", None)
        if sh is not None:  # Lees-Edwards: the shear and its state (degree_oblique, Shear_rate_eff)
            extra["shear"] = np.array([sh[0], float(sh[1]), sh[2], *self.shear_state()])
        np.savez(path, zeta0=z0, zeta1=z1, uk_dc=uk, particles=self.get_particles(),
                 ts=np.int64(t.ts), time=np.array([t.time, t.dt_fluid, t.hdt_fluid, t.dt_md, t.hdt_md]),
                 grid=np.array([c.nx, c.ny, c.nz]), precision=np.int64(c.precision), **extra)

    def load_state(self, path):
        """Restore a state written by save_state (Force_restore_parameters +
        the particle/fluid reads, resume.cxx:507-587, 758-923). The grid,
        particle count and precision must match this run."""
        with np.load(path, allow_pickle=False) as f:
            c = self.params.c
            if tuple(f["grid"]) != (c.nx, c.ny, c.nz) or len(f["particles"]) != self.np:
                raise KhError(f"restart state is for grid {tuple(f['grid'])} with {len(f['particles'])} "
                              f"particles, this run has {(c.nx, c.ny, c.nz)} with {self.np}")
            if int(f["precision"]) != c.precision:
                raise KhError("restart state was written by a run of another precision")
            self.set_particles(f["particles"])
            self.set_fluid(f["zeta0"], f["zeta1"], f["uk_dc"])
            tm = f["time"]
            self.set_time(CTime(int(f["ts"]), *(float(v) for v in tm)))
            if "shear" in f.files:
                rate, ac, freq, deg, sre = (float(v) for v in f["shear"])
                self.set_shear(rate, ac=ac != 0.0, frequency=freq)
                self.set_shear_state(deg, sre)

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

    def set_trajectory(self, nsteps):
        """Record (x, v) of every particle after each of the next nsteps steps on the device."""
        _check(lib().kh_set_trajectory(self.h, nsteps))
        self._traj_steps = nsteps

    def get_trajectory(self, nsteps=None):
        """(nsteps, np, 6) array: x[3], v[3] per particle per recorded step."""
        n = self._traj_steps if nsteps is None else nsteps
        out = np.zeros((n, self.np, 6))
        _check(lib().kh_get_trajectory(self.h, out.ctypes.data, n))
        return out

    def set_shear(self, rate, ac=False, frequency=0.0):
        """Lees-Edwards shear from the next step on (Time_evolution_hydro_OBL):
        DC at `rate`, or AC rate*cos(2 pi frequency t)."""
        _check(lib().kh_set_shear(self.h, float(rate), int(bool(ac)), float(frequency)))
        self._shear = (float(rate), bool(ac), float(frequency))

    def shear_state(self):
        """(degree_oblique, Shear_rate_eff) of a sheared run."""
        g, s = C.c_double(), C.c_double()
        _check(lib().kh_get_shear_state(self.h, C.byref(g), C.byref(s)))
        return g.value, s.value

    def set_shear_state(self, degree_oblique, shear_rate_eff):
        _check(lib().kh_set_shear_state(self.h, float(degree_oblique), float(shear_rate_eff)))

    def shear_record(self):
        """The last sheared step's stress record as a dict (SHEAR_RECORD keys)."""
        o = np.zeros(len(SHEAR_RECORD))
        _check(lib().kh_get_shear_record(self.h, o.ctypes.data))
        return dict(zip(SHEAR_RECORD, o))

    def set_shear_history(self, nsteps):
        """Keep the stress records of the next nsteps steps on the device."""
        _check(lib().kh_set_shear_history(self.h, int(nsteps)))
        self._shear_hist = int(nsteps)

    def get_shear_history(self, nsteps=None):
        """(nsteps, 12) array of stress records (columns as SHEAR_RECORD)."""
        n = getattr(self, "_shear_hist", 0) if nsteps is None else int(nsteps)
        out = np.zeros((n, len(SHEAR_RECORD)))
        _check(lib().kh_get_shear_history(self.h, out.ctypes.data, n))
        return out

    def set_output(self, nframes, every):
        """Record Output_udf frames on the device at the next nframes step
        counts ts with ts % every == 0 (including the current state when it
        qualifies); 0 stops."""
        _check(lib().kh_set_output(self.h, int(nframes), int(every)))
        self._frames = int(nframes)

    def get_output(self, nframes=None):
        """(t, particles): t is (nframes,) step counts, particles an
        (nframes, np) array of FRAME_PARTICLE_DTYPE records."""
        n = getattr(self, "_frames", 0) if nframes is None else int(nframes)
        words = 1 + 34 * self.np
        raw = np.zeros((n, words))
        _check(lib().kh_get_output(self.h, raw.ctypes.data, n))
        t = raw[:, 0].astype(np.int64)
        parts = np.ascontiguousarray(raw[:, 1:]).view(FRAME_PARTICLE_DTYPE).reshape(n, self.np)
        return t, parts

    def idle(self):
        """True when every enqueued step has completed."""
        return _check(lib().kh_query(self.h)) == 1

    def sync(self):
        _check(lib().kh_sync(self.h))

    def set_rigid(self, cum, free=None, v0=None, w0=None):
        """Rigid composite bodies (kh_set_rigid): body k is made of the
        particles cum[k] .. cum[k+1]-1 (their x_nopbc unwrapped); free[k] 1
        "free" / 0 "fix"; v0, w0 (nb, 3) initial body velocities."""
        cum = np.ascontiguousarray(cum, dtype=np.int32)
        fr = None if free is None else np.ascontiguousarray(free, dtype=np.int32)
        v = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        w = None if w0 is None else np.ascontiguousarray(w0, dtype=np.float64)
        _check(lib().kh_set_rigid(self.h, len(cum) - 1, cum.ctypes.data, None if fr is None else fr.ctypes.data,
                                  None if v is None else v.ctypes.data, None if w is None else w.ctypes.data))
        self.nb = len(cum) - 1

    def get_rigid(self):
        """The bodies' state (RIGID_DTYPE records)."""
        out = np.zeros(self.nb, dtype=RIGID_DTYPE)
        _check(lib().kh_get_rigid(self.h, out.ctypes.data, self.nb))
        return out

    def slip_iterations(self):
        """(total slip iterations, slip loops) run so far (squirmer runs)."""
        it, lo = C.c_longlong(0), C.c_longlong(0)
        _check(lib().kh_get_slip_iterations(self.h, C.byref(it), C.byref(lo)))
        return it.value, lo.value

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


def run_bytes(P, nparticles):
    """Device bytes one run of these parameters allocates."""
    return lib().kh_run_bytes(C.byref(P.c), nparticles)


class Server:
    """Continuous batching of independent runs on one GPU (kh_server_*).

    submit() returns an id at once; the run is admitted when a slot and HBM
    budget are free, stays on the device until its steps are done, and its
    per-step trajectory (nsteps, np, 6) lands in the array submit() returned.
    Runs of one size class share batches of `batch_slots` (one launch per
    fluid pass for the whole batch; < 2: every run steps alone).
    """

    def __init__(self, max_active=16, chunk_steps=8, hbm_fraction=0.8, hbm_budget=0, batch_slots=16):
        cfg = ServerConfig(max_active, chunk_steps, hbm_fraction, hbm_budget, batch_slots)
        h = C.c_void_p()
        _check(lib().kh_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, P, particles, nsteps, frame_every=0):
        """Queue a run; returns (id, trajectory array). With frame_every > 0
        the run's Output_udf frames every frame_every steps (ts = 0 to
        nsteps) are kept too: frames(id) returns them once it completes."""
        parts = np.ascontiguousarray(particles, dtype=PARTICLE_DTYPE)
        traj = np.zeros((nsteps, len(parts), 6))
        rid = C.c_longlong()
        if frame_every > 0:
            raw = np.zeros((nsteps // frame_every + 1, 1 + 34 * len(parts)))
            _check(lib().kh_server_submit_output(self.h, C.byref(P.c), parts.ctypes.data, len(parts), nsteps,
                                                 traj.ctypes.data, int(frame_every), raw.ctypes.data,
                                                 C.byref(rid)))
        else:
            raw = None
            _check(lib().kh_server_submit(self.h, C.byref(P.c), parts.ctypes.data, len(parts), nsteps,
                                          traj.ctypes.data, C.byref(rid)))
        self._keep[rid.value] = (P, parts, traj, raw)  # alive until the run completes
        return rid.value, traj

    def frames(self, rid):
        """(t, particles) of a completed run submitted with frame_every:
        t the step counts, particles (nframes, np) FRAME_PARTICLE_DTYPE."""
        raw = self._keep[rid][3]
        if raw is None:
            raise KhError(f"run {rid} was submitted without frames")
        if not self.poll(rid):
            raise KhError(f"run {rid} has not completed")
        np_ = len(self._keep[rid][1])
        t = raw[:, 0].astype(np.int64)
        return t, np.ascontiguousarray(raw[:, 1:]).view(FRAME_PARTICLE_DTYPE).reshape(len(t), np_)

    def poll(self, rid):
        return _check(lib().kh_server_poll(self.h, rid)) == 1

    def wait(self, rid, timeout=600.0):
        if _check(lib().kh_server_wait(self.h, rid, timeout)) != 1:
            raise KhError(f"run {rid} not done after {timeout} s")

    def run_times(self, rid):
        q, r = C.c_double(), C.c_double()
        _check(lib().kh_server_run_times(self.h, rid, C.byref(q), C.byref(r)))
        return q.value, r.value

    def stats(self):
        st = ServerStats()
        _check(lib().kh_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            _check(lib().kh_server_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
