"""Benchmark of the SPM step: steps/s at 256^3 with 2000 colloids, FP64 and
mixed (FP32 fields, FP64 particles), on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints ONE JSON line. `value` is the FP64 steps/s of the whole step (the
reference's own precision; inputs resident in HBM, K steps between two
device synchronisations); the mixed-precision step is timed the same way and
reported beside it under "mixed". Per-kernel times come from HIP events
recorded around every launch on the run's own stream during a second set of
K steps; `roofline` describes the FP64 kernel with the largest share of the
step. `cpu_baseline` times the oracle (the FP64 CPU restatement, OpenMP) on
this host for a bounded sample of the same workload.
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


def step_bytes(n, p, npts, sf, sp):
    """Algorithmic bytes of one reference-order step: 113 full-field passes
    plus 13 accesses per particle point (no squirmer slip)."""
    return sf * field_words(n) * 113 + sp * p * npts * 13


def kernel_bytes(name, n, p, npts, sf, nphi=None, occ=None):
    """Algorithmic HBM bytes of one launch: every distinct input read once
    and every output written once, over the modes and points the kernel's
    function needs (2/3-kept modes |kx|,|ky| < T, kz < T with T = (N+2)/3;
    Sekibun points of every particle). sf = bytes of one field scalar; nphi
    = grid points where phi_sum > 0 (the only points where u* is read and
    f = u_p - phi u* is nonzero; all points when not given). occ: the step
    keeps the per-row occupancy bits (the FP64 plain step; default sf == 8),
    so z_c2r and z_force read 8 bytes per row instead of phi_sum, and
    z_force's clearing of u_p and phi_sum is counted as its writes."""
    t = (n + 2) // 3
    kx, ky, kz = 2 * t - 1, 2 * t - 1, t
    cs = 2 * sf                      # one complex
    kept = kx * ky * kz              # kept modes
    xcol = n * ky * kz               # x lines of the kept (ky, kz) columns
    ycol = n * n * kz                # y/z lines of the kept kz planes
    real = n ** 3
    pt = p * npts
    nphi = real if nphi is None else min(nphi, real)
    occ = (sf == 8) if occ is None else occ
    bits = 8 * n * n * ((n // 4 + 63) // 64)
    table = {
        # fused pipeline
        "x_zeta2u_inv": cs * (2 * kept + 3 * xcol),
        "y_inv": cs * (3 * xcol + 3 * ycol),
        "z_adv": cs * (3 * ycol + 5 * ycol),
        "y_fwd_adv": cs * (5 * ycol + 5 * xcol),
        "x_fwd_adv": cs * (5 * xcol + 5 * kept),
        "adv_ns": cs * (5 + 2 + 2) * kept,
        "x_adv_ns": cs * (5 * xcol + 2 * kept + 2 * kept),
        # u* is read only where phi_sum > 0 (F_H gather, rigidity force):
        # three spectra in, phi_sum everywhere, u* written at those points
        "z_c2r": cs * 3 * ycol + (bits if occ else sf * real) + sf * 3 * nphi,
        # phi_sum everywhere and f only where phi_sum > 0 (elsewhere f = 0):
        # the particle step leaves f = u_p - min(phi_sum,1) u* in the u_p
        # field, reading u* at the points its force gather already read
        "z_force": (bits + 4 * sf * nphi if occ else sf * real) + 3 * sf * nphi + cs * 3 * ycol,
        "y_fwd_force": cs * (3 * ycol + 3 * xcol),
        "x_f_zeta": cs * (3 * xcol + 2 * kept + 2 * kept),
        "fill": sf * real,
        # particle kernels, per Sekibun point of every particle, the accesses
        # of the reference's loops (the kernels' own point cache is not
        # counted): phi_sum read-modify-write (2); force gather, u x 3 and
        # phi_sum (4); u_p scatter, phi_sum and 3 read-modify-writes (7)
        "phi_scatter": 2 * sf * pt,
        "f_hydro": 4 * sf * pt,
        "up_scatter": 7 * sf * pt,
        "particle_step": 11 * sf * pt,
    }
    return table.get(name)


def phi_points(x, n, a, hxi):
    """Grid points (unit spacing, periodic n^3) where some sphere's profile
    is nonzero: minimum-image distance below a + xi/2 (phi(r) > 0 exactly
    there, profile.h:30-50)."""
    rc = a + hxi
    h = int(np.ceil(rc)) + 1
    o = np.arange(-h, h + 1)
    off = np.stack(np.meshgrid(o, o, o, indexing="ij"), -1).reshape(-1, 3)
    base = np.floor(x).astype(np.int64)
    hit = np.zeros(n ** 3, dtype=bool)
    for i in range(0, len(x), 256):
        g = base[i:i + 256, None, :] + off[None, :, :]            # candidate points
        d = g - x[i:i + 256, None, :]
        d -= np.rint(d / n) * n
        inside = (d * d).sum(-1) < rc * rc
        gi = np.mod(g[inside], n)
        hit[(gi[:, 0] * n + gi[:, 1]) * n + gi[:, 2]] = True
    return int(hit.sum())


def time_gpu(precision, n, nparticles, steps, warmup, graph=True, seed=3):
    """K steps between two device synchronisations (the timed region; steps
    after the first replay one captured HIP graph unless graph=False), then
    K more steps uncaptured with a HIP event pair around every launch on the
    run's stream for the per-kernel table."""
    import kapsel3_amd
    from kapsel3_amd import workloads
    P, parts = workloads.sedimentation(n, nparticles, kbt=1e-3, precision=precision, seed=seed)
    run = kapsel3_amd.Run(P, nparticles)
    run.set_particles(parts)
    run.init_fluid()
    run.set_graph(graph)
    run.step(warmup)
    run.sync()
    t0 = time.perf_counter()
    run.step(steps)
    run.sync()
    t1 = time.perf_counter()
    run.timing(True)
    run.timing_reset()
    t2 = time.perf_counter()
    run.step(steps)
    run.sync()
    t3 = time.perf_counter()
    stats = run.timing_stats()
    run.timing(False)
    p = run.get_particles()
    finite = bool(np.isfinite(p["x"]).all() and np.isfinite(p["v"]).all())
    run.close()
    return (t1 - t0), stats, finite, (t3 - t2), (parts, p)


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


def accuracy(mixed, fp64, steps, a):
    """Mixed against FP64 after the same steps from the same start: max
    errors relative to the FP64 velocity scale, displacement and omega scale."""
    p0, pm = mixed
    _, pd = fp64
    vs = max(np.abs(pd["v"]).max(), 1e-300)
    disp = max(np.abs(pd["x_nopbc"] - p0["x_nopbc"]).max(), 1e-300)
    return {"steps": steps, "velocity_rel_err": float(np.abs(pm["v"] - pd["v"]).max() / vs),
            "position_rel_err": float(np.abs(pm["x_nopbc"] - pd["x_nopbc"]).max() / disp),
            "omega_rel_err": float(np.abs(pm["omega"] - pd["omega"]).max() / (vs / a))}


def squirmer_bench(steps, n=256, count=1000):
    """1000 squirmers at 256^3 (B1 = 0.01, B2/B1 uniform in [-2, 2], random
    orientations) at 256^3, steps/s of the whole step with the slip
    iteration, mixed and FP64 (after 10 warm-up steps), and the mean number
    of slip iterations per step in the timed steps."""
    import kapsel3_amd
    from kapsel3_amd import workloads
    from kapsel3_amd._abi import FP64, MIXED
    out = {"workload": f"{count} squirmers a=5 xi=2 at {n}^3, B1 0.01, slip tol 0.005, at most 4 iterations",
           "steps": steps}
    for name, prec in (("mixed", MIXED), ("fp64", FP64)):
        P, parts = workloads.squirmers(n, count, precision=prec)
        run = kapsel3_amd.Run(P, count)
        run.set_particles(parts)
        run.init_fluid()
        run.step(10)  # the Euler step, then both captured graphs (8 steps and 1) are built
        run.sync()
        it0, lo0 = run.slip_iterations()
        t0 = time.perf_counter()
        run.step(steps)
        run.sync()
        t = time.perf_counter() - t0
        it1, lo1 = run.slip_iterations()
        p = run.get_particles()
        run.close()
        out[name] = {"value": steps / t, "unit": "steps/s", "ms_per_step": t / steps * 1e3,
                     "mean_slip_iterations": (it1 - it0) / max(lo1 - lo0, 1),
                     "finite": bool(np.isfinite(p["v"]).all())}
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def shear_bench(steps, n=256, nparticles=2000, rate=1e-3):
    """The Lees-Edwards sheared step (kh_set_shear) on the config-3
    suspension, mixed and FP64, after 2 warm-up steps. The sheared step
    forms only zeta's 2/3-kept modes; the rest are formed once when the
    state is read (settle: one full forward transform). `settle_ms` is that
    completion timed on its own; `value` counts the steps alone (no state
    read), `read_every_step` a caller that reads zeta after every step."""
    import kapsel3_amd
    from kapsel3_amd import workloads
    from kapsel3_amd._abi import FP64, MIXED
    out = {"workload": f"{n}^3, {nparticles} spheres a=5 xi=2, LJ + gravity 1e-3 + kT 1e-3, "
                       f"Lees-Edwards shear rate {rate:g}", "steps": steps}
    for name, prec in (("mixed", MIXED), ("fp64", FP64)):
        P, parts = workloads.sedimentation(n, nparticles, kbt=1e-3, precision=prec, seed=3)
        run = kapsel3_amd.Run(P, nparticles)
        run.set_particles(parts)
        run.init_fluid()
        run.set_shear(rate)
        run.step(2)
        run.sync()
        settle = []
        for _ in range(3):  # the completion alone: queue drained first, then kh_sync settles
            run.step(1)
            while not run.idle():
                time.sleep(1e-4)
            t0 = time.perf_counter()
            run.sync()
            settle.append(time.perf_counter() - t0)
        settle_s = float(np.median(settle))
        t0 = time.perf_counter()
        run.step(steps)
        run.sync()
        t = time.perf_counter() - t0 - settle_s  # the timed chunk ends with one completion
        p = run.get_particles()
        run.close()
        out[name] = {"value": steps / t, "unit": "steps/s", "ms_per_step": t / steps * 1e3,
                     "settle_ms": settle_s * 1e3,
                     "read_every_step": {"value": 1.0 / (t / steps + settle_s), "unit": "steps/s"},
                     "finite": bool(np.isfinite(p["v"]).all())}
    out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def output_bench(steps, every, n=256, nparticles=2000):
    """Output_udf frames recorded on the device inside the captured step
    (kh_set_output) at the config-3 workload, mixed: steps/s with a frame
    every `every` steps against the same run without frames, and the one
    batched device-to-host copy of all frames."""
    import kapsel3_amd
    from kapsel3_amd import workloads
    from kapsel3_amd._abi import MIXED
    P, parts = workloads.sedimentation(n, nparticles, kbt=1e-3, precision=MIXED, seed=3)
    run = kapsel3_amd.Run(P, nparticles)
    run.set_particles(parts)
    run.init_fluid()
    run.step(3)
    res = {"workload": f"{n}^3, {nparticles} spheres, mixed", "steps": steps, "every": every}
    for label in ("without", "with"):
        nframes = steps // every
        if label == "with":
            ts = run.get_time().ts
            run.set_output(nframes, every)
            nframes = len(range((ts + every - 1) // every * every, ts + steps + 1, every))
        run.step(1)  # re-capture the graph
        run.sync()
        t0 = time.perf_counter()
        run.step(steps)
        run.sync()
        res[f"{label}_frames_steps_per_s"] = steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    t, fr = run.get_output()
    res["copy_ms"] = (time.perf_counter() - t0) * 1e3
    res["frames"] = int(len(t))
    res["frame_bytes"] = int(fr.nbytes // max(len(t), 1) + 8)
    res["overhead"] = res["without_frames_steps_per_s"] / res["with_frames_steps_per_s"] - 1.0
    run.close()
    return res


def serve_bench(count, rate, nsteps, max_active, seed=2024):
    """Serving workload: `count` independent runs (grid uniform in {32,48,64,96}^3,
    1-64 colloids a=4, gravity 1e-3, mixed precision, `nsteps` steps each,
    per-step trajectories out) arriving as a Poisson process at `rate`/s,
    served by one continuous-batching Server. rate = None submits every run
    at once (the loop's capacity: runs/s at saturation)."""
    import kapsel3_amd
    from kapsel3_amd import serving
    reqs = serving.requests(count, seed=seed)
    times = serving.arrivals(count, rate, seed=seed + 1) if rate else [0.0] * count
    with kapsel3_amd.Server(max_active=max_active, chunk_steps=8) as srv:
        t0 = time.perf_counter()
        serving.serve(srv, reqs, times, nsteps)
        wall = time.perf_counter() - t0
        st = srv.stats()
    return {"runs": count, "steps_per_run": nsteps, "arrival_rate_per_s": rate, "max_active": max_active,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"], "max_ms": st["max_ms"],
            "completed": st["completed"], "failed": st["failed"], "peak_active": st["peak_active"],
            "peak_bytes": st["peak_bytes"], "wall_s": wall, "runs_per_s": st["runs_per_s"],
            "steps_per_s": st["steps_per_s"], "grid_point_steps_per_s": st["grid_point_steps_per_s"],
            "workload": "grid uniform {32,48,64,96}^3, colloids uniform 1-64 (capped at 20% volume), a=4 xi=2, "
                        "gravity 1e-3, mixed precision, " + ("Poisson arrivals" if rate else "all submitted at once")}


def roofline_of(stats, n, nparticles, npts, sf, pass_ms_per_step, steps, traffic=None, nphi=None,
                traffic_source=None):
    """Roofline entry of the kernel with the largest total time."""
    best = max(stats.items(), key=lambda kv: kv[1][1])
    name, (launches, ms) = best
    per_launch_ms = ms / launches
    b = kernel_bytes(name, n, nparticles, npts, sf, nphi)
    achieved = (b / (per_launch_ms * 1e-3) / 1e9) if b else None
    share = ms / (pass_ms_per_step * steps)
    t = (traffic or {}).get(name)
    return {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": t,
            "traffic_source": traffic_source if t is not None else None,
            "launches": launches, "avg_us": per_launch_ms * 1e3, "bytes_per_launch": b,
            "share_of_step": share}


def kernel_table(stats, n, nparticles, npts, sf, nphi=None):
    out = {}
    for k, (c, ms) in sorted(stats.items()):
        b = kernel_bytes(k, n, nparticles, npts, sf, nphi)
        us = ms / c * 1e3
        out[k] = {"launches": c, "avg_us": us, "bytes_per_launch": b,
                  "gbs": (b / (us * 1e-6) / 1e9) if b else None}
    return out


def load_traffic(path):
    """Measured HBM bytes per launch by kernel (profiler FETCH_SIZE/WRITE_SIZE
    scaled by a plain copy's reading), as written next to the profiles."""
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def copy_ceiling(mib=256, reps=10):
    """A plain device-to-device copy of `mib` MiB (torch), timed with HIP
    events: the bandwidth a pure streaming kernel reaches on this part (read
    plus write bytes / time). Working set 2 x mib MiB, beyond the
    last-level cache."""
    import torch
    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    torch.cuda.empty_cache()
    return {"bytes": 2 * mib * (1 << 20), "avg_us": ms * 1e3, "GBps": 2 * mib * (1 << 20) / (ms * 1e-3) / 1e9,
            "frac_of_peak": 2 * mib * (1 << 20) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--particles", type=int, default=2000)
    ap.add_argument("--mixed-steps", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="time the FP64 step only")
    ap.add_argument("--mixed-only", action="store_true",
                    help="time the mixed step only (for profiling it; value is then the mixed steps/s)")
    ap.add_argument("--no-graph", action="store_true", help="run the timed steps uncaptured")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-squirm", action="store_true", help="skip the squirmer suspension")
    ap.add_argument("--squirm-steps", type=int, default=16)
    ap.add_argument("--no-shear", action="store_true", help="skip the Lees-Edwards sheared suspension")
    ap.add_argument("--shear-steps", type=int, default=20)
    ap.add_argument("--shear-only", action="store_true",
                    help="run only the Lees-Edwards figure and print it (for profiling that step)")
    ap.add_argument("--no-output", action="store_true", help="skip the Output_udf frame recording figure")
    ap.add_argument("--output-steps", type=int, default=50)
    ap.add_argument("--output-every", type=int, default=5)
    ap.add_argument("--serve-runs", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--serve-active", type=int, default=16)
    ap.add_argument("--saturate-runs", type=int, default=256, help="runs submitted at once for the capacity figure")
    ap.add_argument("--saturate-active", type=int, default=32)
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic_256_2000.json"),
                    help="measured HBM bytes per launch (JSON by precision and kernel)")
    ap.add_argument("--no-copy-ceiling", action="store_true", help="skip the 256 MiB device copy measurement")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: one simulation runs on one GPU; --gpus {a.gpus} is not supported "
              "(independent runs shard across GPUs by launching one process per GPU)", file=sys.stderr)
        sys.exit(2)
    from kapsel3_amd._abi import FP64, MIXED
    n, P = a.grid, a.particles
    ceiling = None if a.no_copy_ceiling else copy_ceiling()
    npts = 1935  # Sekibun points of a = 5, xi = 2 (profile.cxx:22-65)
    if a.shear_only:
        print(json.dumps(shear_bench(a.shear_steps, n, P)))
        return

    traffic = load_traffic(a.traffic) if (n, P) == (256, 2000) else None
    tsrc = ("stored: profiles/traffic_256_2000.json, rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this "
            "kernel scaled by a plain copy's reading (not measured in this run)")

    def measure(prec, name, sf, k):
        t, st, ok, tk, state = time_gpu(prec, n, P, k, a.warmup, graph=not a.no_graph)
        ms, kms = t / k * 1e3, tk / k * 1e3
        nphi = phi_points(state[1]["x"], n, 5.0, 1.0) if P else 0
        table = kernel_table(st, n, P, npts, sf, nphi)
        pruned = sum(v["bytes_per_launch"] * v["launches"] / k for v in table.values() if v["bytes_per_launch"])
        return {"value": k / t, "unit": "steps/s", "ms_per_step": ms, "steps": k, "finite": ok,
                "step_roofline": {
                    "pipeline_step_bytes": pruned,
                    "achieved": pruned / (ms * 1e-3) / 1e9, "frac": pruned / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac_of_copy_ceiling": (pruned / (ms * 1e-3) / 1e9 / ceiling["GBps"]) if ceiling else None,
                    "phi_points": nphi,
                    "reference_step_bytes_label": step_bytes(n, P, npts, sf, sf),
                    "note": "pipeline_step_bytes: the algorithmic bytes of every launch of this pruned, fused step "
                            "(kernels table); reference_step_bytes_label: the reference-order step's bytes (113 "
                            "field passes + 13 accesses per Sekibun point), a label only, not a rate"},
                "roofline": roofline_of(st, n, P, npts, sf, kms, k, (traffic or {}).get(name), nphi, tsrc),
                "kernel_pass_ms_per_step": kms, "kernels": table, "_state": state}

    if a.mixed_only:
        head, prec, dt = measure(MIXED, "mixed", 4, a.steps), "mixed", "f32 fields / f64 particles"
    else:
        head, prec, dt = measure(FP64, "fp64", 8, a.steps), "fp64", "f64"
    out = {"metric": METRIC, "value": head["value"], "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dt, "data": "synthetic",
           "config": {"workload": f"SPM step, {n}^3 grid, {P} spheres a=5 xi=2, LJ 12:6 + gravity 1e-3 + kT 1e-3, "
                                  f"dt 0.05, {prec} precision", "grid": n, "particles": P, "precision": prec},
           "roofline": head["roofline"], "cpu_baseline": None,
           "grid_point_steps_per_s": n ** 3 * head["value"], "finite": head["finite"], "graph": not a.no_graph,
           "step_roofline": head["step_roofline"], "copy_ceiling": ceiling,
           "kernel_pass_ms_per_step": head["kernel_pass_ms_per_step"], "kernels": head["kernels"]}
    head_state = head.pop("_state")
    if not a.no_mixed and not a.mixed_only:
        km = a.mixed_steps or a.steps
        mx = measure(MIXED, "mixed", 4, km)
        mx_state = mx.pop("_state")
        mx["grid_point_steps_per_s"] = n ** 3 * mx["value"]
        out["mixed"] = mx
        out["speedup_mixed_vs_fp64"] = mx["value"] / out["value"]
        if km == a.steps:  # same trajectory length: mixed against FP64
            out["accuracy_mixed_vs_fp64"] = accuracy(mx_state, head_state, a.warmup + 2 * a.steps, 5.0)
    if not a.no_squirm:
        out["squirmers"] = squirmer_bench(a.squirm_steps)
    if not a.no_shear:
        out["sheared"] = shear_bench(a.shear_steps)
    if not a.no_output:
        out["output_frames"] = output_bench(a.output_steps, a.output_every)

    if not a.no_serve:
        out["serving"] = serve_bench(a.serve_runs, a.serve_rate, a.serve_steps, a.serve_active)
        if a.saturate_runs > 0:
            out["serving_saturation"] = serve_bench(a.saturate_runs, None, a.serve_steps, a.saturate_active)

    if not a.no_cpu_baseline:
        sps, cores, secs = cpu_baseline(n, P, a.cpu_steps)
        out["cpu_baseline"] = {"value": sps, "unit": "steps/s", "cores": cores, "kind": "port",
                               "sample": f"oracle FP64 step, same {n}^3/{P}-sphere workload, {a.cpu_steps} steps "
                                         f"after 1 warm-up step ({secs:.1f} s)"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
