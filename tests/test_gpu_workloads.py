"""The benchmark configurations at their stated sizes.

- sedimentation (128^3, 200 spheres, LJ + gravity, FP64): the GPU against the
  oracle over 20 steps, <= 1e-10 of each quantity's scale (the oracle's
  1000-step run takes minutes on the host; 20 steps cover Euler, AB2 and
  contacts);
- thermal sedimentation (256^3, 2000 spheres, kT = 1e-3): mixed against FP64 on the GPU
  over the full 1000 steps, particle velocity within 1e-4 of the velocity
  scale (the north-star accuracy bar), position within 1e-4 of the largest
  displacement;
- squirmers (256^3, 1000 squirmers): mixed against FP64 over 100 steps.
"""
import numpy as np
import pytest

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64, MIXED

from test_gpu_step import cmp_particles, rel

pytestmark = pytest.mark.gpu


def test_sedimentation_128_fp64_vs_oracle(gpu):
    print("sedimentation 128^3 x 200, 20 steps vs oracle", flush=True)
    P, parts = workloads.sedimentation(128, 200, seed=2)
    o = kref.Run(P, 200)
    g = gpu.Run(P, 200)
    for r in (o, g):
        r.set_particles(parts)
        r.init_fluid()
        r.step(20)
    cmp_particles(g.get_particles(), o.get_particles(), 1e-10, p0=parts)
    zo, zg = o.get_fluid(), g.get_fluid()
    assert rel(zg[0], zo[0]) < 1e-10 and rel(zg[1], zo[1]) < 1e-10
    o.close()
    g.close()


def _gpu_run(gpu, P, parts, steps):
    r = gpu.Run(P, len(parts))
    r.set_particles(parts)
    r.init_fluid()
    r.step(steps)
    p = r.get_particles()
    r.close()
    return p


def test_thermal_256_mixed_vs_fp64_1000_steps(gpu):
    print("thermal sedimentation 256^3 x 2000, kT 1e-3, 1000 steps mixed vs fp64", flush=True)
    out = {}
    for prec in (FP64, MIXED):
        P, parts = workloads.sedimentation(256, 2000, kbt=1e-3, precision=prec, seed=3)
        out[prec] = _gpu_run(gpu, P, parts, 1000)
    p64, p32 = out[FP64], out[MIXED]
    vs = np.abs(p64["v"]).max()
    ev = np.abs(p32["v"] - p64["v"]).max() / vs
    disp = np.abs(p64["x_nopbc"] - parts["x_nopbc"]).max()
    ex = np.abs(p32["x_nopbc"] - p64["x_nopbc"]).max() / disp
    print(f"velocity err {ev:.2e}, position err {ex:.2e}")
    assert np.isfinite(p32["v"]).all() and ev < 1e-4 and ex < 1e-4


def test_squirmers_256_mixed_vs_fp64(gpu):
    print("squirmers 256^3 x 1000, 100 steps mixed vs fp64", flush=True)
    out = {}
    for prec in (FP64, MIXED):
        P, parts = workloads.squirmers(256, 1000, precision=prec)
        out[prec] = _gpu_run(gpu, P, parts, 100)
    p64, p32 = out[FP64], out[MIXED]
    vs = np.abs(p64["v"]).max()
    ev = np.abs(p32["v"] - p64["v"]).max() / vs
    ew = np.abs(p32["omega"] - p64["omega"]).max() / (vs / 5.0)
    print(f"velocity err {ev:.2e}, omega err {ew:.2e}")
    assert vs > 1e-3 and ev < 1e-4 and ew < 1e-4
