"""Squirmer slip (SW_JANUS_SLIP, sp_3d_ns.cxx:121-161) on the GPU against
the oracle: each slip function alone, then whole steps on both step paths
(FP64, <= 1e-10 of each quantity's scale), and the mixed path against the
FP64 GPU path (<= 1e-4)."""
import numpy as np
import pytest

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64, JAX_Y, JAX_Z, JP_NONE, JP_SLIP, MIXED, Params, make_species

from test_gpu_step import cmp_particles, rel

pytestmark = pytest.mark.gpu


def squirmer_system(n=32, precision=FP64, iters=4, tol=0.005, scale=1.0):
    sp = [make_species(mass_ratio=1.0, propulsion=JP_SLIP, janus_axis=JAX_Z, slip_vel=0.01, slip_mode=1.3,
                       rotlet_c1=2e-3, rotlet_c2=-1e-3),
          make_species(mass_ratio=1.0, propulsion=JP_SLIP, janus_axis=JAX_Y, slip_vel=0.01, slip_mode=-0.8),
          make_species(mass_ratio=1.5, propulsion=JP_NONE)]
    P = Params(n, a=4.0, a_xi=2.0, dt=0.05, species=sp, rotation=1, g=0.0, precision=precision,
               slip_iter_max=iters, slip_tol=tol)
    pos = np.array([[6.3, 7.6, 8.2], [20.1, 22.4, 9.7], [12.5, 24.0, 23.3], [26.0, 9.1, 24.8]]) * scale
    parts = workloads.particles_at(pos, spec=np.array([0, 1, 2, 0]), q=workloads.random_quaternions(4, 3))
    return P, parts


def both(P, parts):
    import kapsel3_amd
    o = kref.Run(P, len(parts))
    g = kapsel3_amd.Run(P, len(parts))
    o.set_particles(parts)
    g.set_particles(parts)
    return o, g


MOMENTS = ["mass", "mass_center", "inertia", "surface_mass", "surface_mass_center", "surface_inertia",
           "surface_dv", "surface_dw"]


def test_slip_functions_fp64(gpu):
    print("slip functions 32^3 x 4 (3 squirmers)", flush=True)
    P, parts = squirmer_system()
    o, g = both(P, parts)
    o.init_fluid()
    g.init_fluid()
    o.apply("zeta_k2u")
    g.apply("zeta_k2u")
    o.apply("slip_momentum")
    g.apply("slip_momentum")
    po, pg = o.get_particles(), g.get_particles()
    for k in MOMENTS:
        s = np.abs(po[k]).max()
        assert np.abs(pg[k] - po[k]).max() <= 1e-12 * max(s, 1e-300), k
    o.apply("slip_scatter")
    g.apply("slip_scatter")
    for d in range(3):
        assert rel(g.get_field(f"up{d}"), o.get_field(f"up{d}")) < 1e-12
    o.apply("make_phi")
    g.apply("make_phi")
    o.apply("f_hydro")
    g.apply("f_hydro")
    o.apply("slip_iteration")
    g.apply("slip_iteration")
    cmp_particles(g.get_particles(), o.get_particles(), 1e-10, keys=["v", "omega", "v_slip", "omega_slip"], a=4.0)
    for d in range(3):
        assert rel(g.get_field(f"u{d}"), o.get_field(f"u{d}")) < 1e-10
    o.close()
    g.close()


# 32^3: the fused step's slip loop decided on the device (captured as
# graphs); 48^3: the same step with the loop decided on the host (NZ not a
# power of two)
@pytest.mark.parametrize("path,n", [("fused", 32), ("reference", 32), ("fused", 48)])
def test_squirmer_steps_fp64(gpu, path, n):
    print(f"squirmers {n}^3 x 4, 40 steps, {path}", flush=True)
    P, parts = squirmer_system(n=n, scale=n / 32)
    o, g = both(P, parts)
    g.set_step_path(path)
    o.init_fluid()
    g.init_fluid()
    o.step(40)
    g.step(40)
    it, loops = g.slip_iterations()
    assert loops == 40 and 40 <= it <= 4 * 40, (it, loops)
    po, pg = o.get_particles(), g.get_particles()
    cmp_particles(pg, po, 1e-10, keys=["x", "x_nopbc", "v", "omega", "q", "v_slip", "omega_slip"], p0=parts, a=4.0)
    zo, zg = o.get_fluid(), g.get_fluid()
    assert rel(zg[0], zo[0]) < 1e-10 and rel(zg[1], zo[1]) < 1e-10
    assert np.abs(po["v"][:, :]).max() > 1e-4  # the squirmers swim
    o.close()
    g.close()


def test_squirmers_mixed_vs_fp64(gpu):
    print("squirmers mixed vs fp64, 64^3 x 12, 60 steps", flush=True)
    import kapsel3_amd
    out = {}
    for prec in (FP64, MIXED):
        P, parts = workloads.squirmers(64, 12, a=5.0, precision=prec, seed=7)
        g = kapsel3_amd.Run(P, len(parts))
        g.set_particles(parts)
        g.init_fluid()
        g.step(60)
        out[prec] = g.get_particles()
        g.close()
    p64, p32 = out[FP64], out[MIXED]
    vs = np.abs(p64["v"]).max()
    assert vs > 1e-4
    assert np.abs(p32["v"] - p64["v"]).max() / vs < 1e-4
    assert np.abs(p32["omega"] - p64["omega"]).max() / (vs / 5.0) < 1e-4
