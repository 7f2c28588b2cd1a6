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


def _frames_close(t, fr, t_ref, fr_ref, tol):
    assert list(t) == list(t_ref)
    for k in range(len(t)):
        for a, _ in PAIRS:
            ref = fr_ref[k][a]
            scale = max(np.abs(fr_ref[a]).max(), 1e-300)
            assert np.abs(fr[k][a] - ref).max() <= tol * scale, (t[k], a)
        assert not fr[k]["torque_r"].any()


def test_frames_against_oracle(gpu):
    """kh_set_output frames of an FP64 run (graph-captured steps) against the
    oracle's particle state at the same step counts."""
    from test_output_udf import oracle_frames
    print("output frames 32^3 x 6 thermal vs oracle", flush=True)
    P, parts = workloads.sedimentation(32, 6, a=4, kbt=1e-3, gap=0.4, seed=11, precision=FP64)
    t_ref, fr_ref = oracle_frames(P, parts, every=4, nframes=6)
    r = gpu.Run(P, len(parts))
    r.set_particles(parts)
    r.init_fluid()
    r.set_output(6, 4)
    r.step(21)
    t, fr = r.get_output()
    r.close()
    _frames_close(t, fr, t_ref, fr_ref, 1e-10)


def test_served_frames_against_oracle_and_udf_file(gpu, tmp_path):
    """A served FP64 run (kh_server_submit_output) delivers Output_udf frames
    every 5 steps; they match the oracle's state at those steps, and written
    as UDF records (output.cxx:268-375) they read back unchanged."""
    from kapsel3_amd import output
    from test_output_udf import oracle_frames
    print("served frames 48^3 x 10 vs oracle", flush=True)
    P, parts = workloads.sedimentation(48, 10, a=4, kbt=0.0, gap=1.0, seed=5, precision=FP64)
    with gpu.Server(max_active=2, chunk_steps=8) as srv:
        rid, traj = srv.submit(P, parts, 30, frame_every=5)
        rid2, _ = srv.submit(P, parts, 12)           # a second run in the mix, without frames
        srv.wait(rid)
        srv.wait(rid2)
        t, fr = srv.frames(rid)
    t_ref, fr_ref = oracle_frames(P, parts, every=5, nframes=7)
    _frames_close(t, fr, t_ref, fr_ref, 1e-10)
    # the last frame is the state after the last step, as the trajectory's
    assert np.array_equal(fr[-1]["R"], traj[-1, :, :3]) and np.array_equal(fr[-1]["v"], traj[-1, :, 3:])
    path = tmp_path / "served.udf"
    output.write_output_udf(path, t, fr)
    t2, fr2 = output.read_output_udf(path)
    assert np.array_equal(t2, t) and np.array_equal(fr2, fr)
