"""Rigid composite bodies on the GPU (kh_set_rigid) against the oracle
through the C ABI: the initial block (mass, centre, principal moments and
frame of each body, the beads' orientation and velocity) and whole steps of
a bent trimer, a dimer and a one-bead body in LJ contact under gravity, on
the fused and the reference-order path, with the reference's LJ torque
bookkeeping and without it."""
import numpy as np
import pytest

from oracle import kref
from kapsel3_amd._abi import FP64, MIXED, Params, make_species, new_particles

from test_gpu_step import cmp_particles, rel

pytestmark = pytest.mark.gpu

N = 32
# bent trimer across the x boundary, a dimer in LJ contact with it, a lone bead
BEADS = np.array([[20.5, 15.2, 16.3], [27.8, 15.9, 16.1], [21.4, 22.7, 17.0],
                  [4.4, 15.0, 16.0], [11.9, 15.6, 16.2], [6.2, 5.1, 26.3]])
CUM = [0, 3, 5, 6]
V0 = np.array([[0.01, -0.004, 0.006], [-0.008, 0.007, 0.002], [0.0, 0.003, -0.005]])
W0 = np.array([[0.002, -0.001, 0.003], [-0.002, 0.003, 0.001], [0.001, 0.0, 0.002]])


def system(compat=1, precision=FP64):
    sp = [make_species(mass_ratio=1.5), make_species(mass_ratio=2.0)]
    P = Params(N, a=4.0, a_xi=2.0, dt=0.05, kbt=0.0, g=1e-3, g_dir=2, rotation=1, fix_cell=(1, 1, 1),
               lj_truncate=1, lj_cell_compat=compat, species=sp, precision=precision)
    p = new_particles(len(BEADS))
    p["x_nopbc"] = BEADS
    p["x"] = np.mod(BEADS, N)
    p["spec"] = [0, 0, 0, 1, 1, 0]
    return P, p


def both(compat=1):
    P, p = system(compat)
    o, g = kref.Run(P, len(p)), pytest.importorskip("kapsel3_amd").Run(P, len(p))
    for r in (o, g):
        r.set_particles(p)
        r.set_rigid(CUM, free=[1, 1, 1], v0=V0, w0=W0)
    return o, g


def cmp_bodies(bg, bo, tol):
    vs = np.abs(bo["vG"]).max()
    ws = np.abs(bo["omegaG"]).max()
    assert np.abs(bg["mass"] - bo["mass"]).max() <= 1e-13 * bo["mass"].max()
    assert np.abs(bg["Ibody"] - bo["Ibody"]).max() <= 1e-12 * bo["Ibody"].max()
    assert np.abs(bg["xG_nopbc"] - bo["xG_nopbc"]).max() <= tol * N
    assert np.abs(bg["vG"] - bo["vG"]).max() <= tol * vs
    assert np.abs(bg["omegaG"] - bo["omegaG"]).max() <= tol * ws
    assert np.abs(bg["q"] - bo["q"]).max() <= tol
    fs = max(np.abs(bo["f_hydro"]).max(), 1e-300)
    assert np.abs(bg["f_hydro"] - bo["f_hydro"]).max() <= tol * fs
    assert np.abs(bg["torque_hydro"] - bo["torque_hydro"]).max() <= tol * fs * 10
    rs = max(np.abs(bo["f_r"]).max(), 1e-300)
    assert np.abs(bg["f_r"] - bo["f_r"]).max() <= tol * rs
    assert np.abs(bg["torque_r"] - bo["torque_r"]).max() <= tol * rs * 10


def test_rigid_initial_block(gpu):
    print("rigid initial block: trimer, dimer, one bead at 32^3", flush=True)
    o, g = both()
    bo, bg = o.get_rigid(), g.get_rigid()
    cmp_bodies(bg, bo, 1e-12)
    po, pg = o.get_particles(), g.get_particles()
    for k in ("q", "v", "omega", "x", "x_previous"):
        assert np.abs(pg[k] - po[k]).max() <= 1e-12 * max(np.abs(po[k]).max(), 1.0), k


@pytest.mark.parametrize("path", ["fused", "reference"])
@pytest.mark.parametrize("compat", [1, 0])
def test_rigid_steps_against_oracle(gpu, path, compat):
    print(f"rigid bodies 32^3, 40 steps, path {path}, lj compat {compat}", flush=True)
    o, g = both(compat)
    g.set_step_path(path)
    p0 = o.get_particles()
    for r in (o, g):
        r.init_fluid()
        r.step(40)
    bo, bg = o.get_rigid(), g.get_rigid()
    assert np.abs(bo["f_r"]).max() > 0          # the bodies touch
    cmp_bodies(bg, bo, 1e-10)
    pg, po = g.get_particles(), o.get_particles()
    cmp_particles(pg, po, 1e-10,
                  ["v", "omega", "q", "q_old", "v_old", "omega_old", "fr_previous", "f_hydro_previous"], a=4.0)
    # bead positions against the beads' own displacement (the rigid position
    # update places beads by x only and leaves x_nopbc as it was,
    # update_Particle_Configuration, rigid.h:168-185)
    d0 = po["x"] - p0["x"]
    d0 -= np.rint(d0 / N) * N
    dx = pg["x"] - po["x"]
    dx -= np.rint(dx / N) * N
    assert np.abs(d0).max() > 1e-3
    assert np.abs(dx).max() <= 1e-10 * np.abs(d0).max()
    assert np.array_equal(pg["x_nopbc"], po["x_nopbc"])
    zo, zg = o.get_fluid(), g.get_fluid()
    assert rel(zg[0], zo[0]) < 1e-10 and rel(zg[1], zo[1]) < 1e-10


def test_rigid_mixed_precision_follows_fp64(gpu):
    """FP32 fields with FP64 body sums and integration: within 1e-5 of the
    FP64 GPU run after 40 steps."""
    import kapsel3_amd
    print("rigid bodies 32^3, 40 steps, mixed vs FP64", flush=True)
    out = []
    for prec in (FP64, MIXED):
        P, p = system(precision=prec)
        r = kapsel3_amd.Run(P, len(p))
        r.set_particles(p)
        r.set_rigid(CUM, v0=V0, w0=W0)
        r.init_fluid()
        r.step(40)
        out.append(r.get_rigid())
    b64, b32 = out
    assert np.abs(b32["vG"] - b64["vG"]).max() < 1e-5 * np.abs(b64["vG"]).max()
    assert np.abs(b32["omegaG"] - b64["omegaG"]).max() < 1e-5 * np.abs(b64["omegaG"]).max()
