"""Output_udf frames recorded on the device (output.cxx:268-375): each frame
must be an exact copy of the particle state at its step count, taken at
ts % every == 0 before the step of that ts (sp_3d_ns.cxx:863-950), with the
previous step's forces and torques as the reference writes them. Checked
against kh_get_particles snapshots of the same run, captured-graph steps
included."""
import numpy as np
import pytest

from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64, MIXED

from test_gpu_slip import squirmer_system

pytestmark = pytest.mark.gpu

PAIRS = [("R", "x"), ("R_raw", "x_nopbc"), ("v", "v"), ("q", "q"), ("omega", "omega"),
         ("f_hydro", "f_hydro_previous"), ("torque_hydro", "torque_hydro_previous"), ("f_r", "fr_previous"),
         ("f_slip", "f_slip_previous"), ("torque_slip", "torque_slip_previous")]


def _frames_match_snapshots(gpu, P, parts, every, nframes, start):
    r = gpu.Run(P, len(parts))
    r.set_particles(parts)
    r.init_fluid()
    r.step(start)
    r.set_output(nframes, every)
    snaps, ts = {}, start
    while len(snaps) < nframes:
        if ts % every == 0:
            snaps[ts] = r.get_particles()
        r.step(1)
        ts += 1
    t, fr = r.get_output()
    assert list(t) == sorted(snaps)
    for k, s in enumerate(sorted(snaps)):
        for a, b in PAIRS:
            assert np.array_equal(fr[k][a], snaps[s][b]), (s, a)
        assert not fr[k]["torque_r"].any()
    r.close()


@pytest.mark.parametrize("precision", [FP64, MIXED], ids=["fp64", "mixed"])
def test_frames_thermal_sedimentation(gpu, precision):
    print(f"output frames 32^3 x 6 thermal, precision {precision}", flush=True)
    P, parts = workloads.sedimentation(32, 6, a=4, kbt=1e-3, gap=0.4, seed=11, precision=precision)
    _frames_match_snapshots(gpu, P, parts, every=3, nframes=4, start=0)   # frames at 0, 3, 6, 9
    _frames_match_snapshots(gpu, P, parts, every=4, nframes=3, start=2)   # frames at 4, 8, 12


def test_frames_squirmers(gpu):
    print("output frames 32^3 x 4 squirmers", flush=True)
    P, parts = squirmer_system()
    _frames_match_snapshots(gpu, P, parts, every=2, nframes=3, start=1)   # frames at 2, 4, 6
