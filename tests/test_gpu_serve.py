"""Continuous batching (kh_server_*): every run of a mixed stream completes
on the GPU, its trajectory matches the same run served alone (the runs share
the device, not state), admission respects the HBM budget, and a run's
recorded trajectory equals the states kh_step produces."""
import numpy as np
import pytest

from oracle import kref
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
    """The serving workload's sizes (32, 48, 64 and 96^3, a = 4, up to 64
    colloids, gravity) for its 200 steps: 24 runs in one mix (stepped in
    batches per size class), three of them (one per size class up to 96^3)
    again alone, outside any batch."""
    print("served mix of 24 runs (32-96^3) x 200 steps vs alone", flush=True)
    reqs = serving.requests(24, seed=11, sizes=(32, 48, 64, 96), pmax=64, precision=MIXED)
    with gpu.Server(max_active=8, chunk_steps=8) as srv:
        mix = [srv.submit(P, parts, 200) for P, parts in reqs]
        for rid, _ in mix:
            srv.wait(rid)
        st = srv.stats()
    assert st["completed"] == 24 and st["failed"] == 0 and st["peak_active"] > 1
    sizes = [P.c.nx for P, _ in reqs]
    picks = [sizes.index(n) for n in (32, 64, 96)]
    for k in picks:  # the same runs, each alone on a fresh server
        P, parts = reqs[k]
        print(f"  alone: run {k}, {P.c.nx}^3 x {len(parts)}", flush=True)
        with gpu.Server(max_active=1, batch_slots=0) as srv:
            rid, alone = srv.submit(P, parts, 200)
            srv.wait(rid)
        got = mix[k][1]
        vs = max(np.abs(alone[..., 3:]).max(), 1e-12)
        disp = max(np.abs(alone[..., :3] - parts["x"][None]).max(), 1e-12)
        assert np.abs(got[..., 3:] - alone[..., 3:]).max() <= 1e-5 * vs
        assert np.abs(got[..., :3] - alone[..., :3]).max() <= 1e-5 * disp


def test_served_fp64_run_per_size_matches_oracle(gpu):
    """One FP64 run of each serving size (32, 48, 64, 96^3; a = 4, xi = 2,
    gravity 1e-3, seeded non-overlapping colloids) served together for 200
    steps: every step's (x, v) against the oracle's trajectory of the same
    run (Time_evolution_hydro, sp_3d_ns.cxx:41-194, looped as 857-956)."""
    runs = []
    for n, npart, seed in ((32, 24, 21), (48, 40, 22), (64, 64, 23), (96, 64, 24)):
        P, parts = workloads.sedimentation(n, npart, a=4.0, kbt=0.0, precision=FP64, seed=seed, gap=1.0)
        runs.append((P, parts))
    print("served FP64 runs 32/48/64/96^3 x 200 steps", flush=True)
    with gpu.Server(max_active=4, chunk_steps=8) as srv:
        out = [srv.submit(P, parts, 200) for P, parts in runs]
        for rid, _ in out:
            srv.wait(rid)
    for (P, parts), (_, traj) in zip(runs, out):
        n = P.c.nx
        print(f"  oracle {n}^3 x {len(parts)}", flush=True)
        o = kref.Run(P, len(parts))
        o.set_particles(parts)
        o.init_fluid()
        ref = np.zeros_like(traj)
        for s in range(200):
            o.step(1)
            p = o.get_particles()
            ref[s, :, :3], ref[s, :, 3:] = p["x"], p["v"]
        o.close()
        vs = np.abs(ref[..., 3:]).max()
        d = ref[..., :3] - parts["x"][None]
        d -= np.rint(d / n) * n
        disp = np.abs(d).max()
        dx = traj[..., :3] - ref[..., :3]
        dx -= np.rint(dx / n) * n
        assert vs > 0 and disp > 0
        assert np.abs(traj[..., 3:] - ref[..., 3:]).max() < 1e-10 * vs, n
        assert np.abs(dx).max() < 1e-10 * disp, n


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
