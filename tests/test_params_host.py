"""Parameter checks of kh_run_create and the footprint of kh_run_bytes, on
the host: both run before any device call, so invalid input is refused
with a clear error even where no GPU is present."""
import pytest

import kapsel3_amd
from kapsel3_amd._abi import FP64, MIXED, Params, make_species


def _refused(P, np_, match):
    with pytest.raises(kapsel3_amd.KhError, match=match):
        kapsel3_amd.Run(P, np_)


@pytest.mark.parametrize("n", [7, 10, 1024, (32, 32, 1000)])
def test_unsupported_grid_sizes(n):
    _refused(Params(n), 0, "not a supported FFT length")


def test_nonpositive_physics():
    _refused(Params(32, dt=0.0), 1, "must be positive")
    _refused(Params(32, eta=-1.0), 1, "must be positive")
    _refused(Params(32, species=[make_species(mass_ratio=0.0)]), 1, "mass_ratio must be positive")


def test_particle_domain_must_fit_the_grid():
    # a = 4, xi = 2: Sekibun half-width 8 does not fit an 8-point side
    _refused(Params(8, a=4.0, a_xi=2.0), 1, "does not fit the grid")


def test_other_options():
    _refused(Params(32, lj_powers=3), 1, "LJ_powers")
    _refused(Params(32, g_dir=3), 1, "g_dir")
    _refused(Params(32), -1, "negative particle count")


def test_run_bytes():
    lib = kapsel3_amd.lib()
    import ctypes as C
    b64 = lib.kh_run_bytes(C.byref(Params(256, precision=FP64).c), 2000)
    b32 = lib.kh_run_bytes(C.byref(Params(256, precision=MIXED).c), 2000)
    assert 2.0e9 < b64 < 4.0e9 and 1.0e9 < b32 < 2.0e9        # a small share of 288 GB
    assert 1.8 < b64 / b32 < 2.1
    assert kapsel3_amd.run_bytes(Params(64), 10) < kapsel3_amd.run_bytes(Params(64), 100)
    assert lib.kh_run_bytes(None, 1) < 0
