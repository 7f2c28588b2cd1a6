"""Captured-graph steps against uncaptured steps of the same run.

Steps with ts >= 1 replay a captured HIP graph by default (one launch per
8 steps when at least 8 are asked for, else one per step). Each replay is
checked against the uncaptured step, with host transfers (eager work and
synchronisations) between the replays, after a change of dt (which must
re-capture) and after kh_apply (which must run one uncaptured step).
Atomic scatters make the summation order vary between runs, so the bar is
1e-12 of the field scale (FP64), not bit equality."""
import numpy as np
import pytest

from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64

pytestmark = pytest.mark.gpu


def _pair(gpu, n=32, np_=6, kbt=1e-3):
    P, parts = workloads.sedimentation(n, np_, a=4, kbt=kbt, gap=0.4, seed=11, precision=FP64)
    runs = []
    for graph in (True, False):
        r = gpu.Run(P, np_)
        r.set_particles(parts)
        r.init_fluid()
        r.set_graph(graph)
        runs.append(r)
    return runs


def _close(a, b, tol=1e-12):
    za, zb = a.get_fluid(), b.get_fluid()
    pa, pb = a.get_particles(), b.get_particles()
    for x, y in zip(za[:2], zb[:2]):
        s = max(np.abs(y).max(), 1e-300)
        assert np.abs(x - y).max() <= tol * s
    vs = max(np.abs(pb["v"]).max(), 1e-300)
    assert np.abs(pa["v"] - pb["v"]).max() <= tol * vs
    assert np.abs(pa["x_nopbc"] - pb["x_nopbc"]).max() <= 1e-12 * max(np.abs(pb["x_nopbc"]).max(), 1.0)
    assert a.get_time().ts == b.get_time().ts


def test_graph_replays_match_uncaptured(gpu):
    print("graph replays vs uncaptured, 32^3 x 6", flush=True)
    g, e = _pair(gpu)
    for k in range(6):  # first step uncaptured (Euler), then replays
        g.step(1)
        e.step(1)
        _close(g, e)
    g.step(5)
    e.step(5)
    _close(g, e)
    g.step(17)  # two replays of the 8-step graph, then one single step
    e.step(17)
    _close(g, e)
    g.close()
    e.close()


def test_graph_recaptures_on_dt_change_and_after_apply(gpu):
    print("graph re-capture on dt change / apply", flush=True)
    g, e = _pair(gpu)
    g.step(3)
    e.step(3)
    for r in (g, e):
        t = r.get_time()
        t.dt_fluid *= 0.5
        t.hdt_fluid *= 0.5
        t.dt_md *= 0.5
        t.hdt_md *= 0.5
        r.set_time(t)
    g.step(3)
    e.step(3)
    _close(g, e)
    for r in (g, e):
        r.apply("zeta_k2u")
    g.step(3)
    e.step(3)
    _close(g, e)
    g.close()
    e.close()
