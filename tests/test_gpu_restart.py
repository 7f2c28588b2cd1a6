"""Restart: a run saved after some steps and resumed in a fresh kh_run must
follow the uninterrupted run (Save_Restart_udf / Force_restore_parameters,
resume.cxx:10-74, 507-587). The saved state carries the AB2 history in the
Particle records and jikan.ts, which keys the thermal noise, so the resumed
run takes AB2 steps and draws the same noise. Atomic scatters reorder sums
between runs, so the bar is 1e-12 of each quantity's scale for FP64 and
1e-5 for the FP32/FP64 mixed path."""
import numpy as np
import pytest

from kapsel3_amd import workloads
from kapsel3_amd._abi import FP64, MIXED

from test_gpu_slip import squirmer_system

pytestmark = pytest.mark.gpu


def _close(a, b, tol):
    for x, y in zip(a.get_fluid()[:2], b.get_fluid()[:2]):
        assert np.abs(x - y).max() <= tol * max(np.abs(y).max(), 1e-300)
    pa, pb = a.get_particles(), b.get_particles()
    for k in ("v", "omega", "f_hydro", "v_old"):
        assert np.abs(pa[k] - pb[k]).max() <= tol * max(np.abs(pb[k]).max(), 1e-300), k
    assert np.abs(pa["x_nopbc"] - pb["x_nopbc"]).max() <= 1e-12 * max(np.abs(pb["x_nopbc"]).max(), 1.0)
    assert a.get_time().ts == b.get_time().ts


def _resume(gpu, P, parts, tmp_path, before, after, tol, shear=None):
    a = gpu.Run(P, len(parts))
    a.set_particles(parts)
    a.init_fluid()
    if shear is not None:
        a.set_shear(*shear)
    a.step(before)
    path = tmp_path / "state.npz"
    a.save_state(path)
    a.step(after)
    b = gpu.Run(P, len(parts))
    b.load_state(path)
    assert b.get_time().ts == before
    b.step(after)
    _close(b, a, tol)
    if shear is not None:
        assert b.shear_state() == a.shear_state()
    a.close()
    b.close()


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-12), (MIXED, 1e-5)], ids=["fp64", "mixed"])
def test_resume_thermal_sedimentation(gpu, tmp_path, precision, tol):
    print(f"restart 32^3 x 6 thermal, precision {precision}", flush=True)
    P, parts = workloads.sedimentation(32, 6, a=4, kbt=1e-3, gap=0.4, seed=11, precision=precision)
    _resume(gpu, P, parts, tmp_path, before=4, after=5, tol=tol)


@pytest.mark.parametrize("ac", [False, True], ids=["dc", "ac"])
def test_resume_sheared(gpu, tmp_path, ac):
    """The Lees-Edwards state (rate, strain, Shear_rate_eff) travels with the
    npz state: the resumed run continues exactly."""
    print(f"restart 32^3 x 6 sheared, ac={ac}", flush=True)
    P, parts = workloads.sedimentation(32, 6, a=4, kbt=1e-3, gap=0.4, seed=11, precision=FP64)
    _resume(gpu, P, parts, tmp_path, before=4, after=5, tol=1e-12, shear=(0.05, ac, 0.2))


def test_resume_squirmers_fp64(gpu, tmp_path):
    print("restart 32^3 x 4 squirmers", flush=True)
    P, parts = squirmer_system()
    _resume(gpu, P, parts, tmp_path, before=3, after=4, tol=1e-12)


def test_resume_rejects_other_grid(gpu, tmp_path):
    print("restart grid mismatch", flush=True)
    P, parts = workloads.sedimentation(32, 6, a=4, gap=0.4, seed=11)
    a = gpu.Run(P, len(parts))
    a.set_particles(parts)
    a.init_fluid()
    a.save_state(tmp_path / "s.npz")
    a.close()
    Q, parts_q = workloads.sedimentation(48, 6, a=4, gap=0.4, seed=11)
    b = gpu.Run(Q, len(parts_q))
    with pytest.raises(gpu.KhError):
        b.load_state(tmp_path / "s.npz")
    b.close()
