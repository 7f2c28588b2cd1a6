"""Seeded inputs for the benchmark configurations of the SPM step.

Each builder returns (Params, particles) in KAPSEL's own structures; the
physics follows the reference decks (Examples/02/gravity.udf for
sedimentation, input.cxx for units): DX = 1, RHO = 1, ETA = 1, manual dt.
Positions are drawn by random sequential addition from a numpy seed with a
minimum surface gap, since the reference's own placement (glibc rand()
seeded 10, init_particle.cxx:18, 65-83) is not portable.
"""
import numpy as np

from ._abi import (FP64, JAX_Z, JP_MOTOR, JP_NONE, JP_SLIP, MIXED, Params, make_species, new_particles)


def place_spheres(n, box, a, gap=1.0, seed=0, max_tries=200000):
    """Random sequential addition in a periodic box: centres at least
    2a + gap apart (minimum image). Returns an (n, 3) array; raises if the
    packing cannot be reached within max_tries draws."""
    rng = np.random.default_rng(seed)
    box = np.asarray(box, dtype=np.float64)
    dmin2 = (2.0 * a + gap) ** 2
    pos = np.zeros((n, 3))
    k = tries = 0
    while k < n:
        if tries >= max_tries:
            raise ValueError(f"could only place {k} of {n} spheres")
        tries += 1
        c = rng.random(3) * box
        if k:
            d = pos[:k] - c
            d -= np.rint(d / box) * box
            if np.min(np.einsum("ij,ij->i", d, d)) < dmin2:
                continue
        pos[k] = c
        k += 1
    return pos


def particles_at(pos, spec=None, q=None):
    p = new_particles(len(pos))
    p["x"] = pos
    p["x_nopbc"] = pos
    p["x_previous"] = pos
    if spec is not None:
        p["spec"] = spec
    if q is not None:
        p["q"] = q
        p["q_old"] = q
    return p


def random_quaternions(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def plain_single(n=32, precision=FP64):
    """One neutral sphere a=4, xi=2 driven by a constant body force
    1e-3 along +z (a motor force, since gravity on a neutral sphere is zero,
    md_force.cxx:314), fluid at rest, dt 0.05, FIX_CELL on all axes."""
    sp = make_species(mass_ratio=1.0, propulsion=JP_MOTOR, janus_axis=JAX_Z, janus_force=(0.0, 0.0, 1e-3))
    P = Params(n, a=4.0, a_xi=2.0, dt=0.05, species=[sp], rotation=1, fix_cell=(1, 1, 1), lj_truncate=1,
               g=0.0, precision=precision)
    c = n / 2.0
    return P, particles_at(np.array([[c, c, c]]))


def sedimentation(n=128, nparticles=200, a=5.0, kbt=0.0, precision=FP64, seed=1, mass_ratio=2.0, gap=1.0):
    """Sedimentation: spheres a=5, xi=2, LJ 12:6 repulsive (cutoff
    2^(1/6) 2a), eps=1, gravity 1e-3 along -z, rho_p/rho = 2, fluid and
    particles at rest, dt 0.05. kbt > 0 adds the thermostat (alpha = 1)."""
    sp = make_species(mass_ratio=mass_ratio, propulsion=JP_NONE)
    P = Params(n, a=a, a_xi=2.0, dt=0.05, species=[sp], rotation=1, fix_cell=(1, 1, 1), lj_truncate=1,
               lj_powers=0, epsilon=1.0, g=1e-3, g_dir=2, kbt=kbt, precision=precision, seed=seed)
    pos = place_spheres(nparticles, (n, n, n), a, gap=gap, seed=seed)
    return P, particles_at(pos)


def squirmers(n=256, nparticles=1000, a=5.0, b1=0.01, precision=MIXED, seed=5, slip_iter=4, slip_tol=0.005):
    """Squirmers with B1 = 0.01 and per-particle B2/B1 uniform in
    [-2, 2] (one species per particle, as the reference keeps slip
    parameters per species), random orientations, no gravity, kT = 0."""
    rng = np.random.default_rng(seed)
    beta = rng.uniform(-2.0, 2.0, nparticles)
    species = [make_species(mass_ratio=1.0, propulsion=JP_SLIP, janus_axis=JAX_Z, slip_vel=b1, slip_mode=b)
               for b in beta]
    P = Params(n, a=a, a_xi=2.0, dt=0.05, species=species, rotation=1, fix_cell=(1, 1, 1), lj_truncate=1,
               g=0.0, precision=precision, seed=seed, slip_iter_max=slip_iter, slip_tol=slip_tol)
    pos = place_spheres(nparticles, (n, n, n), a, gap=1.0, seed=seed)
    return P, particles_at(pos, spec=np.arange(nparticles), q=random_quaternions(nparticles, seed + 1))


def hasimoto_velocity(force, a, eta, box_volume, nparticles=1):
    """Sedimentation speed of a simple-cubic periodic array (Hasimoto 1959):
    U/U_Stokes = 1 - 1.7601 c^(1/3) + c - 1.5593 c^2, c = volume fraction."""
    c = nparticles * 4.0 / 3.0 * np.pi * a ** 3 / box_volume
    return force / (6 * np.pi * eta * a) * (1 - 1.7601 * c ** (1 / 3) + c - 1.5593 * c ** 2)
