"""The serving workload: a seeded stream of independent runs of different
box sizes and colloid counts, submitted at Poisson arrival times to one
`Server` (continuous batching on one GPU)."""
import time

import numpy as np

from . import workloads
from ._abi import MIXED

SIZES = (32, 48, 64, 96)


def max_colloids(n, a=4.0, fraction=0.2):
    """Colloids a random sequential placement still fits: volume fraction
    capped at `fraction` (24 at 32^3 with a = 4, gap 1)."""
    return max(1, int(fraction * n ** 3 / (4.0 / 3.0 * np.pi * a ** 3)))


def requests(count=512, seed=2024, sizes=SIZES, pmax=64, a=4.0, precision=MIXED):
    """(arrival offset s, Params, particles) for `count` runs: grid uniform
    over `sizes`, colloid count uniform in 1..pmax (capped by what fits),
    a = 4, xi = 2, gravity 1e-3, seeded non-overlapping positions (min gap 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        npart = min(int(rng.integers(1, pmax + 1)), max_colloids(n, a))
        P, parts = workloads.sedimentation(n, npart, a=a, kbt=0.0, precision=precision,
                                           seed=int(rng.integers(1 << 30)), gap=1.0)
        out.append((P, parts))
    return out


def arrivals(count, rate, seed=2025):
    """Poisson process: cumulative exponential inter-arrival times (s)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, count))


def serve(server, reqs, times, nsteps=200):
    """Submit each request at its arrival time, wait for all; returns the
    run ids and their trajectory arrays."""
    t0 = time.perf_counter()
    ids = []
    for (P, parts), t in zip(reqs, times):
        dt = t - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids.append(server.submit(P, parts, nsteps))
    for rid, _ in ids:
        server.wait(rid)
    return ids
