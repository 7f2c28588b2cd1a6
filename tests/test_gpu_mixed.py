"""Mixed precision (FP32 fields, FP64 particles) against the FP64 GPU step,
which is itself pinned to the oracle. Tolerance: particle velocity within
1e-4 of the velocity scale, positions within 1e-4 of the displacement."""
import numpy as np
import pytest

from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64, MIXED

pytestmark = pytest.mark.gpu


def _run(gpu, P, parts, nsteps):
    g = gpu.Run(P, len(parts))
    g.set_particles(parts)
    g.init_fluid()
    g.step(nsteps)
    p = g.get_particles()
    z = g.get_fluid()
    g.close()
    return p, z


@pytest.mark.parametrize("n,np_,kbt,steps", [(64, 24, 1e-3, 200), (48, 10, 0.0, 200), (256, 300, 1e-3, 40)])
def test_mixed_vs_fp64(gpu, n, np_, kbt, steps):
    print(f"mixed vs fp64 {n}^3 x {np_} kT={kbt}", flush=True)
    P64, parts = workloads.sedimentation(n, np_, kbt=kbt, gap=0.5, seed=6, precision=FP64)
    P32, _ = workloads.sedimentation(n, np_, kbt=kbt, gap=0.5, seed=6, precision=MIXED)
    p64, z64 = _run(gpu, P64, parts, steps)
    p32, z32 = _run(gpu, P32, parts, steps)
    vs = np.abs(p64["v"]).max()
    assert vs > 0
    ev = np.abs(p32["v"] - p64["v"]).max() / vs
    disp = np.abs(p64["x_nopbc"] - parts["x_nopbc"]).max()
    ex = np.abs(p32["x_nopbc"] - p64["x_nopbc"]).max() / disp
    ew = np.abs(p32["omega"] - p64["omega"]).max() / (vs / 5.0)
    print(f"velocity err {ev:.2e}  position err {ex:.2e}  omega err {ew:.2e}")
    assert ev < 1e-4 and ex < 1e-4 and ew < 1e-4
    assert np.isfinite(z32[0]).all()
