"""kh_A2a_k_noalloc, the forward transform that allocates no device memory
(its staging is a fixed device array), against kh_A2a_k on a run of the
same grid, through the C ABI: two calls in a row, the second on a smaller
grid, each bit-identical to kh_A2a_k and within 1e-13 of numpy's rfftn;
grids beyond the fixed array are refused before any GPU work."""
import numpy as np
import pytest

from kapsel3_amd._abi import Params

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    nx, ny, nz = shape
    a = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    a[:, :, :nz] = np.random.default_rng(seed).standard_normal((nx, ny, nz))
    return a


def test_noalloc_matches_run_transform(gpu):
    # the second call is smaller: the part of the array it uses still holds
    # the first call's data
    for shape, seed in (((64, 48, 96), 1), ((16, 32, 24), 2)):
        print(f"kh_A2a_k_noalloc {shape} against kh_A2a_k", flush=True)
        a = _field(shape, seed)
        got = gpu.A2a_k_noalloc(a)
        with gpu.Run(Params(shape), 0) as run:
            ref = run.A2a_k(a)
        assert np.array_equal(got, ref)
        nz = shape[2]
        want = np.fft.rfftn(a[:, :, :nz])
        gc = got[:, :, 0::2] + 1j * got[:, :, 1::2]
        assert np.abs(gc - want).max() / np.abs(want).max() < 1e-13


def test_noalloc_refuses_grids_beyond_its_array(gpu):
    import kapsel3_amd
    print("kh_A2a_k_noalloc size and length checks", flush=True)
    with pytest.raises(kapsel3_amd.KhError, match="scratch"):
        gpu.A2a_k_noalloc(np.zeros((256, 256, 258)))
    with pytest.raises(kapsel3_amd.KhError):
        gpu.A2a_k_noalloc(np.zeros((10, 16, 18)))
