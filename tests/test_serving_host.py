"""Host side of the serving workload (grids 32-96^3, 1-64 spheres): the seeded request stream
and the Poisson arrival process. No GPU."""
import numpy as np

from kapsel3_amd import serving
from kapsel3_amd._abi import MIXED


def test_request_stream_shapes():
    reqs = serving.requests(64, seed=3)
    sizes = [P.c.nx for P, _ in reqs]
    assert set(sizes) <= set(serving.SIZES) and len(set(sizes)) == len(serving.SIZES)
    for P, parts in reqs:
        n = P.c.nx
        assert P.c.nx == P.c.ny == P.c.nz and P.c.precision == MIXED and P.c.a == 4.0
        assert 1 <= len(parts) <= min(64, serving.max_colloids(n))
        x = parts["x"]
        d = x[:, None, :] - x[None, :, :]
        d -= np.rint(d / n) * n
        r = np.sqrt((d ** 2).sum(-1)) + np.eye(len(parts)) * 1e9
        assert r.min() >= 2 * 4.0 + 1.0 - 1e-9  # non-overlapping, gap >= 1
    again = serving.requests(64, seed=3)
    assert all(np.array_equal(a[1]["x"], b[1]["x"]) for a, b in zip(reqs, again))


def test_colloid_cap():
    assert serving.max_colloids(32) == 24 and serving.max_colloids(96) > 64


def test_poisson_arrivals():
    t = serving.arrivals(20000, 20.0, seed=1)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0) and abs(gaps.mean() - 0.05) < 0.002
    assert abs(gaps.std() - 0.05) < 0.003  # exponential: std = mean
