"""Edge cases of the step on the GPU: no particles (a pure fluid whose
decay the slaved-Euler scheme fixes in closed form, and against the
oracle), grids whose three sides differ, and the largest grid the FFT
engines are built for (512^3), where index arithmetic is widest."""
import math

import numpy as np
import pytest

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64, MIXED, Params, make_species

from test_gpu_step import cmp_particles, rel, run_both

pytestmark = pytest.mark.gpu


def spec(a):
    return a[:, :, 0::2] + 1j * a[:, :, 1::2]


def _taylor_green(r, n, A):
    g = 2 * np.pi * np.arange(n) / n
    X, Y = np.meshgrid(g, g, indexing="ij")
    u = [(A * np.sin(X) * np.cos(Y))[:, :, None] * np.ones((1, 1, n)),
         (-A * np.cos(X) * np.sin(Y))[:, :, None] * np.ones((1, 1, n)), np.zeros((n, n, n))]
    for d in range(3):
        f = np.zeros(r.shape)
        f[:, :, :n] = u[d]
        r.set_field(f"u{d}", f)
    r.apply("u2zeta_k")


@pytest.mark.parametrize("precision,path,tol", [(FP64, "fused", 1e-12), (FP64, "reference", 1e-12),
                                                (MIXED, "fused", 1e-5)], ids=["fp64-fused", "fp64-ref", "mixed"])
def test_fluid_only_taylor_green(gpu, precision, path, tol):
    """No particles: u.grad(u) of a Taylor-Green vortex is a gradient, so
    every whole step multiplies zeta by exp(-nu k^2 dt), k^2 = 2 (2pi/N)^2
    (NS_solver_slavedEuler, fluid_solver.cxx:348-374; the particle terms
    vanish with phi = 0)."""
    print(f"taylor-green 32^3, no particles, {path}, precision {precision}", flush=True)
    n, A, eta, dt, steps = 32, 5e-2, 0.8, 0.05, 10
    P = Params(n, eta=eta, dt=dt, precision=precision)
    r = gpu.Run(P, 0)
    r.set_step_path(path)
    _taylor_green(r, n, A)
    z0 = spec(r.get_fluid()[1]).copy()
    r.step(steps)
    z1 = spec(r.get_fluid()[1])
    k2 = 2 * (2 * np.pi / n) ** 2
    m = np.abs(z0) > 1e-6 * np.abs(z0).max()
    assert m.sum() >= 4
    assert np.abs(z1[m] / z0[m] - math.exp(-eta * k2 * dt * steps)).max() < tol
    # and nothing appears outside the vortex's modes
    assert np.abs(z1[~m]).max() < tol * np.abs(z0).max()
    assert r.get_time().ts == steps
    r.close()


def test_fluid_only_vs_oracle(gpu):
    """No particles, a random band-limited start, 20 steps: FP64 GPU against
    the oracle (nonlinear term active)."""
    print("random fluid 32^3, no particles, vs oracle", flush=True)
    n = 32
    P = Params(n, eta=0.5, dt=0.05, fix_cell=(0, 0, 0))
    rng = np.random.default_rng(7)
    u = [rng.standard_normal((n, n, n)) * 1e-2 for _ in range(3)]
    runs = []
    for r in (kref.Run(P, 0), gpu.Run(P, 0)):
        for d in range(3):
            f = np.zeros(r.shape)
            f[:, :, :n] = u[d]
            r.set_field(f"u{d}", f)
        r.apply("u2zeta_k")
        r.step(20)
        runs.append(r)
    zo, zg = runs[0].get_fluid(), runs[1].get_fluid()
    assert rel(zg[0], zo[0]) < 1e-10 and rel(zg[1], zo[1]) < 1e-10
    assert np.abs(zg[2] - zo[2]).max() <= 1e-10 * max(np.abs(zo[2]).max(), 1e-300)


@pytest.mark.parametrize("shape", [(32, 64, 128), (128, 32, 64)])
def test_noncubic_grid_vs_oracle(gpu, shape):
    """Three different sides (x, y, z roles exercised separately), spheres,
    gravity, LJ and noise: 12 steps against the oracle."""
    print(f"non-cubic {shape}, 8 spheres, vs oracle", flush=True)
    Q = Params(shape, a=4.0, a_xi=2.0, dt=0.05, species=[make_species(mass_ratio=2.0)], g=1e-3, g_dir=2,
               kbt=1e-3)
    pos = workloads.place_spheres(8, np.array(shape, dtype=float), 4.0, gap=0.5, seed=9)
    parts = workloads.particles_at(pos)
    o, g = run_both(Q, parts, 12)
    cmp_particles(g.get_particles(), o.get_particles(), 1e-10, p0=parts, a=4.0)
    zo, zg = o.get_fluid(), g.get_fluid()
    assert rel(zg[0], zo[0]) < 1e-10 and rel(zg[1], zo[1]) < 1e-10


def test_largest_grid_mixed_vs_fp64(gpu):
    """512^3 with 400 spheres, 4 steps: the mixed path against the FP64 path
    (the oracle is too slow at this size); both finite, particle velocities
    within 1e-4 of their scale."""
    print("512^3 x 400, mixed vs fp64, 4 steps", flush=True)
    out = []
    for precision in (FP64, MIXED):
        P, parts = workloads.sedimentation(512, 400, kbt=1e-3, precision=precision, seed=12)
        r = gpu.Run(P, len(parts))
        r.set_particles(parts)
        r.init_fluid()
        r.step(4)
        p = r.get_particles()
        assert np.isfinite(p["v"]).all() and np.isfinite(p["x"]).all()
        out.append(p)
        r.close()
    vs = np.abs(out[0]["v"]).max()
    assert vs > 0
    assert np.abs(out[1]["v"] - out[0]["v"]).max() <= 1e-4 * vs
