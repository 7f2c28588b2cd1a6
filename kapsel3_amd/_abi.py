"""ctypes mirrors of the C-ABI structs in include/kapsel_hip.h and
include/kapsel_particle.h (KAPSEL's Particle/CTime, variable.h:18-77)."""
import ctypes as C

import numpy as np

DIM = 3


class CTime(C.Structure):
    _fields_ = [("ts", C.c_int), ("time", C.c_double), ("dt_fluid", C.c_double),
                ("hdt_fluid", C.c_double), ("dt_md", C.c_double), ("hdt_md", C.c_double)]


_D3 = (np.float64, (3,))
_D33 = (np.float64, (3, 3))
# Particle (variable.h:27-77), 1048 bytes, natural alignment
PARTICLE_FIELDS = [
    ("spec", np.int32), ("_pad0", np.int32),
    ("x", _D3), ("x_previous", _D3), ("x_nopbc", _D3),
    ("v", _D3), ("v_old", _D3), ("v_slip", _D3),
    ("f_hydro", _D3), ("f_hydro_previous", _D3), ("f_hydro1", _D3), ("f_slip", _D3), ("f_slip_previous", _D3),
    ("fr", _D3), ("fr_previous", _D3), ("torque_r", _D3), ("torque_r_previous", _D3),
    ("omega", _D3), ("omega_old", _D3), ("omega_slip", _D3),
    ("torque_hydro", _D3), ("torque_hydro_previous", _D3), ("torque_hydro1", _D3), ("torque_slip", _D3),
    ("torque_slip_previous", _D3),
    ("momentum_depend_fr", _D3),
    ("QR", _D33), ("QR_old", _D33),
    ("q", (np.float64, (4,))), ("q_old", (np.float64, (4,))),
    ("mass", np.float64), ("mass_center", _D3), ("inertia", _D33),
    ("surface_mass", np.float64), ("surface_mass_center", _D3), ("surface_inertia", _D33),
    ("surface_dv", _D3), ("surface_dw", _D3),
]
PARTICLE_DTYPE = np.dtype(PARTICLE_FIELDS)
assert PARTICLE_DTYPE.itemsize == 1048
assert PARTICLE_DTYPE.fields["x"][1] == 8 and PARTICLE_DTYPE.fields["v"][1] == 80
assert PARTICLE_DTYPE.fields["f_hydro"][1] == 152 and PARTICLE_DTYPE.fields["omega"][1] == 368
assert PARTICLE_DTYPE.fields["q"][1] == 728

# one rigid body (kh_get_rigid, KH_RIGID_WORDS doubles)
RIGID_DTYPE = np.dtype([("xG", "f8", 3), ("xG_nopbc", "f8", 3), ("vG", "f8", 3), ("omegaG", "f8", 3),
                        ("f_hydro", "f8", 3), ("torque_hydro", "f8", 3), ("f_r", "f8", 3), ("torque_r", "f8", 3),
                        ("mass", "f8"), ("Ibody", "f8", 3), ("q", "f8", 4)])
assert RIGID_DTYPE.itemsize == 32 * 8

# enum JP / JAX (input.h:48-49)
JP_MOTOR, JP_SLIP, JP_OBSTACLE, JP_NONE = 0, 1, 2, 3
JAX_X, JAX_Y, JAX_Z, JAX_NONE = 0, 1, 2, 3
FP64, MIXED = 0, 1

FIELD = {"u0": 0, "u1": 1, "u2": 2, "up0": 3, "up1": 4, "up2": 5, "phi": 6, "phi_sum": 7,
         "f0": 8, "f1": 9, "f2": 10}
FN = {name: i for i, name in enumerate([
    "ns_solver", "zeta_k2u", "u2zeta_k", "position_euler", "position_ab2", "make_phi", "f_hydro",
    "velocity_euler", "velocity_ab2", "thermostat", "make_up", "f_particle_sole", "add_f_particle",
    "slip_momentum", "slip_scatter", "slip_iteration", "obl_ns_solver", "obl_zeta_k2u", "obl_u2zeta_k",
    "obl_oblique2u"])}


class KhSpecies(C.Structure):
    _fields_ = [("mass_ratio", C.c_double), ("propulsion", C.c_int), ("janus_axis", C.c_int),
                ("janus_force", C.c_double * 3), ("janus_torque", C.c_double * 3),
                ("slip_vel", C.c_double), ("slip_mode", C.c_double),
                ("rotlet_c1", C.c_double), ("rotlet_c2", C.c_double)]


class KhParams(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("rho", C.c_double), ("eta", C.c_double), ("kbt", C.c_double),
                ("alpha_v", C.c_double), ("alpha_o", C.c_double), ("dt", C.c_double),
                ("a", C.c_double), ("a_xi", C.c_double), ("g", C.c_double), ("g_dir", C.c_int),
                ("rotation", C.c_int), ("fix_cell", C.c_int * 3), ("lj_truncate", C.c_int),
                ("lj_powers", C.c_int), ("epsilon", C.c_double), ("lj_cell_compat", C.c_int),
                ("nspecies", C.c_int), ("species", C.POINTER(KhSpecies)),
                ("slip_iter_max", C.c_int), ("slip_tol", C.c_double), ("precision", C.c_int),
                ("seed", C.c_uint64), ("stream_id", C.c_uint32)]


def make_species(mass_ratio=1.0, propulsion=JP_NONE, janus_axis=JAX_NONE, janus_force=(0, 0, 0),
                 janus_torque=(0, 0, 0), slip_vel=0.0, slip_mode=0.0, rotlet_c1=0.0, rotlet_c2=0.0):
    s = KhSpecies()
    s.mass_ratio, s.propulsion, s.janus_axis = mass_ratio, propulsion, janus_axis
    s.janus_force[:] = list(janus_force)
    s.janus_torque[:] = list(janus_torque)
    s.slip_vel, s.slip_mode, s.rotlet_c1, s.rotlet_c2 = slip_vel, slip_mode, rotlet_c1, rotlet_c2
    return s


class Params:
    """Owns a KhParams plus the species array it points to."""

    def __init__(self, n, dx=1.0, rho=1.0, eta=1.0, kbt=0.0, alpha_v=1.0, alpha_o=1.0, dt=0.05,
                 a=5.0, a_xi=2.0, g=0.0, g_dir=2, rotation=1, fix_cell=(1, 1, 1), lj_truncate=1,
                 lj_powers=0, epsilon=1.0, lj_cell_compat=1, species=None, slip_iter_max=4,
                 slip_tol=0.005, precision=FP64, seed=12345, stream_id=0):
        nx, ny, nz = (n, n, n) if isinstance(n, int) else n
        species = species or [make_species()]
        self._species = (KhSpecies * len(species))(*species)
        p = KhParams()
        p.nx, p.ny, p.nz = nx, ny, nz
        p.dx, p.rho, p.eta, p.kbt, p.alpha_v, p.alpha_o, p.dt = dx, rho, eta, kbt, alpha_v, alpha_o, dt
        p.a, p.a_xi, p.g, p.g_dir = a, a_xi, g, g_dir
        p.rotation = rotation
        p.fix_cell[:] = list(fix_cell)
        p.lj_truncate, p.lj_powers, p.epsilon, p.lj_cell_compat = lj_truncate, lj_powers, epsilon, lj_cell_compat
        p.nspecies = len(species)
        p.species = C.cast(self._species, C.POINTER(KhSpecies))
        p.slip_iter_max, p.slip_tol, p.precision = slip_iter_max, slip_tol, precision
        p.seed, p.stream_id = seed, stream_id
        self.c = p

    @property
    def shape(self):
        return (self.c.nx, self.c.ny, self.c.nz)

    @property
    def nz_(self):
        return 2 * (self.c.nz // 2 + 1)

    def field_size(self):
        return self.c.nx * self.c.ny * self.nz_


def new_particles(n):
    p = np.zeros(n, dtype=PARTICLE_DTYPE)
    p["q"][:, 0] = 1.0
    p["q_old"][:, 0] = 1.0
    return p


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))
