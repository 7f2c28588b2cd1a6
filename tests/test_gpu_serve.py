"""Continuous batching (kh_server_*): every run of a mixed stream completes
on the GPU, its trajectory matches the same run served alone (the runs share
the device, not state), admission respects the HBM budget, and a run's
recorded trajectory equals the states kh_step produces."""
import numpy as np
import pytest

from kapsel3_amd import serving, workloads
from kapsel3_amd._abi import MIXED, FP64

pytestmark = pytest.mark.gpu


def test_trajectory_record_matches_steps(gpu):
    print("trajectory record 32^3 x 5", flush=True)
    P, parts = workloads.sedimentation(32, 5, a=4.0, precision=FP64, seed=3, gap=1.0)
    r = gpu.Run(P, 5)
    r.set_particles(parts)
    r.init_fluid()
    r.set_trajectory(12)
    xs = []
    for _ in range(12):
        r.step(1)
        p = r.get_particles()
        xs.append(np.concatenate([p["x"], p["v"]], 1))
    tr = r.get_trajectory()
    assert np.array_equal(tr, np.stack(xs))
    r.close()


def test_served_runs_match_alone(gpu):
    print("served mix of 24 runs vs alone", flush=True)
    reqs = serving.requests(24, seed=11, sizes=(32, 48, 64), pmax=12, precision=MIXED)
    with gpu.Server(max_active=8, chunk_steps=4) as srv:
        mix = [srv.submit(P, parts, 40) for P, parts in reqs]
        for rid, _ in mix:
            srv.wait(rid)
        st = srv.stats()
    assert st["completed"] == 24 and st["failed"] == 0 and st["peak_active"] > 1
    for k in (0, 7, 19):  # the same runs, each alone on a fresh server
        P, parts = reqs[k]
        with gpu.Server(max_active=1) as srv:
            rid, alone = srv.submit(P, parts, 40)
            srv.wait(rid)
        got = mix[k][1]
        vs = max(np.abs(alone[..., 3:]).max(), 1e-12)
        disp = max(np.abs(alone[..., :3] - parts["x"][None]).max(), 1e-12)
        assert np.abs(got[..., 3:] - alone[..., 3:]).max() <= 1e-5 * vs
        assert np.abs(got[..., :3] - alone[..., :3]).max() <= 1e-5 * disp


def test_admission_by_budget(gpu):
    print("admission under a two-run budget", flush=True)
    reqs = serving.requests(6, seed=5, sizes=(64,), pmax=4, precision=MIXED)
    need = max(gpu.run_bytes(P, len(p)) for P, p in reqs) + 6 * 8 * 30 * 4
    with gpu.Server(max_active=8, hbm_budget=int(2.2 * need)) as srv:
        ids = [srv.submit(P, parts, 30) for P, parts in reqs]
        for rid, _ in ids:
            srv.wait(rid)
        st = srv.stats()
    assert st["completed"] == 6 and st["peak_active"] <= 2 and st["peak_bytes"] <= st["budget"]
