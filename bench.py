"""Benchmark of the SPM step: steps/s at 256^3 with 2000 colloids, mixed
(FP32 fields, FP64 particles) and FP64, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line. `value` is the mixed-precision steps/s of the whole
step (inputs resident in HBM, K steps between two device synchronisations).
The FP64 step is timed the same way and reported beside it. Per-kernel
times come from HIP events recorded around every launch on the run's own
stream during the timed steps; `roofline` describes the kernel with the
largest share of the step. `cpu_baseline` times the oracle (the FP64 CPU
restatement, OpenMP) on this host for a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "steps/s at 256³, 2000 colloids (FP64 vs mixed); p50/p99 served-run latency"
HBM_PEAK_GBS = 8000.0


def field_words(n):
    """M(N) = N^2 (N+2): words of one padded field (real or half-spectrum)."""
    return n * n * (n + 2)


# Algorithmic field passes per launch (reads of distinct inputs + writes of
# outputs, each counted once at full padded size), by kernel name.
FIELD_PASSES = {
    "fft_z_r2c": 2, "fft_z_c2r": 2, "fft_y_fwd": 2, "fft_y_inv": 2, "fft_x_fwd": 2, "fft_x_inv": 2,
    "zeta_to_u": 5, "products": 8, "adv_ns": 9, "solenoidal": 6, "u_to_zeta": 5, "f_nonsole": 10,
    "add3": 9, "fill": 1,
}
# Particle accesses per Sekibun point per launch
POINT_ACCESSES = {"phi_scatter": 2, "f_hydro": 4, "up_scatter": 7}


def step_bytes(n, p, npts, sf, sp):
    """Algorithmic bytes of one reference step: 113 field passes plus 13
    accesses per particle point (no squirmer slip)."""
    return sf * field_words(n) * 113 + sp * p * npts * 13


def kernel_bytes(name, n, p, npts, sf):
    if name in FIELD_PASSES:
        return FIELD_PASSES[name] * field_words(n) * sf
    if name in POINT_ACCESSES:
        return POINT_ACCESSES[name] * p * npts * sf
    return None


def time_gpu(precision, n, nparticles, steps, warmup, seed=3):
    import kapsel3_amd
    from kapsel3_amd import workloads
    P, parts = workloads.sedimentation(n, nparticles, kbt=1e-3, precision=precision, seed=seed)
    run = kapsel3_amd.Run(P, nparticles)
    run.set_particles(parts)
    run.init_fluid()
    run.step(warmup)
    run.sync()
    run.timing(True)
    run.timing_reset()
    run.sync()
    t0 = time.perf_counter()
    run.step(steps)
    run.sync()
    t1 = time.perf_counter()
    stats = run.timing_stats()
    p = run.get_particles()
    finite = bool(np.isfinite(p["x"]).all() and np.isfinite(p["v"]).all())
    run.close()
    return (t1 - t0), stats, finite


def cpu_baseline(n, nparticles, steps, seed=3):
    from oracle import kref
    from kapsel3_amd import workloads
    from kapsel3_amd._abi import FP64
    P, parts = workloads.sedimentation(n, nparticles, kbt=1e-3, precision=FP64, seed=seed)
    r = kref.Run(P, nparticles)
    r.set_particles(parts)
    r.init_fluid()
    r.step(1)  # warm-up (Euler first step)
    t0 = time.perf_counter()
    r.step(steps)
    t1 = time.perf_counter()
    cores = kref.lib().kr_threads()
    r.close()
    return steps / (t1 - t0), cores, t1 - t0


def roofline_of(stats, n, nparticles, npts, sf, wall_ms_per_step, steps):
    best = max(stats.items(), key=lambda kv: kv[1][1])
    name, (launches, ms) = best
    per_launch_ms = ms / launches
    b = kernel_bytes(name, n, nparticles, npts, sf)
    achieved = (b / (per_launch_ms * 1e-3) / 1e9) if b else None
    share = ms / (wall_ms_per_step * steps)
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
            "launches": launches, "avg_us": per_launch_ms * 1e3, "bytes_per_launch": b,
            "share_of_step": share}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--particles", type=int, default=2000)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (torch) so profiler byte counters can be scaled")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(independent runs shard across GPUs by launching one process per GPU)", file=sys.stderr)
        sys.exit(2)
    from kapsel3_amd._abi import FP64, MIXED
    n, P = a.grid, a.particles
    if a.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    npts = 1935  # Sekibun points of a = 5, xi = 2 (profile.cxx:22-65)

    t_mix, st_mix, ok_mix = time_gpu(MIXED, n, P, a.steps, a.warmup)
    ms_mix = t_mix / a.steps * 1e3
    out = {"metric": METRIC, "value": a.steps / t_mix, "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_mix, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32 fields / f64 particles", "data": "synthetic",
           "config": {"workload": f"SPM step, {n}^3 grid, {P} spheres a=5 xi=2, LJ 12:6 + gravity 1e-3 + kT 1e-3, "
                                  "dt 0.05, mixed precision", "grid": n, "particles": P, "precision": "mixed"},
           "grid_point_steps_per_s": n ** 3 * a.steps / t_mix, "finite": ok_mix}
    sb = step_bytes(n, P, npts, 4, 4)
    out["step_roofline"] = {"algorithmic_bytes": sb, "achieved": sb / (ms_mix * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": sb / (ms_mix * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out["roofline"] = roofline_of(st_mix, n, P, npts, 4, ms_mix, a.steps)
    out["kernels_mixed"] = {k: {"launches": c, "avg_us": ms / c * 1e3} for k, (c, ms) in sorted(st_mix.items())}

    if not a.no_fp64:
        k64 = a.fp64_steps or a.steps
        t64, st64, ok64 = time_gpu(FP64, n, P, k64, a.warmup)
        ms64 = t64 / k64 * 1e3
        sb64 = step_bytes(n, P, npts, 8, 8)
        out["fp64"] = {"value": k64 / t64, "unit": "steps/s", "ms_per_step": ms64, "steps": k64, "finite": ok64,
                       "step_roofline": {"algorithmic_bytes": sb64, "achieved": sb64 / (ms64 * 1e-3) / 1e9,
                                         "frac": sb64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS},
                       "roofline": roofline_of(st64, n, P, npts, 8, ms64, k64)}
        out["speedup_mixed_vs_fp64"] = out["value"] / out["fp64"]["value"]

    if not a.no_cpu_baseline:
        sps, cores, secs = cpu_baseline(n, P, a.cpu_steps)
        out["cpu_baseline"] = {"value": sps, "unit": "steps/s", "cores": cores, "kind": "port",
                               "sample": f"oracle FP64 step, same {n}^3/{P}-sphere workload, {a.cpu_steps} steps "
                                         f"after 1 warm-up step ({secs:.1f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
