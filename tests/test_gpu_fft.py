"""FFT parity on the GPU: kh_A2a_k / kh_A_k2a (A2a_k / A_k2a semantics,
fft_wrapper.h:641-681) against the oracle and numpy, through the C ABI."""
import numpy as np
import pytest

from oracle import kref
from kapsel3_amd._abi import FP64, MIXED, Params

pytestmark = pytest.mark.gpu

SIZES = [(8, 8, 8), (16, 16, 16), (32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128),
         (16, 32, 64), (24, 12, 48), (256, 256, 256)]


def _real_field(shape, seed):
    nx, ny, nz = shape
    a = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    a[:, :, :nz] = np.random.default_rng(seed).standard_normal((nx, ny, nz))
    return a


@pytest.mark.parametrize("shape", SIZES, ids=lambda s: "x".join(map(str, s)))
def test_forward_inverse_fp64(gpu, shape):
    print(f"fft fp64 {shape}", flush=True)
    nx, ny, nz = shape
    a = _real_field(shape, 7)
    with gpu.Run(Params(shape), 0) as run:
        g = run.A2a_k(a)
        o = kref.fft_forward(a)
        err = np.abs(g - o).max() / np.abs(o).max()
        assert err < 1e-13, err
        if nx * ny * nz <= 64 ** 3:
            ref = np.fft.rfftn(a[:, :, :nz])
            gc = g[:, :, 0::2] + 1j * g[:, :, 1::2]
            assert np.abs(gc - ref).max() / np.abs(ref).max() < 1e-13
        back = run.A_k2a(g)
        ob = kref.fft_inverse(o)
        assert np.abs(back[:, :, :nz] - ob[:, :, :nz]).max() < 1e-12
        assert np.abs(back[:, :, :nz] - a[:, :, :nz]).max() < 1e-12


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 48, 48), (128, 128, 128)], ids=lambda s: "x".join(map(str, s)))
def test_forward_inverse_fp32(gpu, shape):
    print(f"fft fp32 {shape}", flush=True)
    nz = shape[2]
    a = _real_field(shape, 8)
    with gpu.Run(Params(shape, precision=MIXED), 0) as run:
        g = run.A2a_k(a)
        o = kref.fft_forward(a)
        assert np.abs(g - o).max() / np.abs(o).max() < 2e-6
        back = run.A_k2a(g)
        assert np.abs(back[:, :, :nz] - a[:, :, :nz]).max() < 2e-5


@pytest.mark.parametrize("precision,tol", [(FP64, 1e-13), (MIXED, 2e-6)], ids=["fp64", "fp32"])
def test_forward_inverse_512(gpu, precision, tol):
    """The largest length the engines are built for (512: a full wavefront
    per z line, one-workgroup-per-CU column tiles)."""
    print(f"fft 512^3 precision {precision}", flush=True)
    shape = (512, 512, 512)
    a = _real_field(shape, 9)
    with gpu.Run(Params(shape, precision=precision), 0) as run:
        g = run.A2a_k(a)
        o = kref.fft_forward(a)
        assert np.abs(g - o).max() / np.abs(o).max() < tol
        back = run.A_k2a(g)
        assert np.abs(back[:, :, :512] - a[:, :, :512]).max() < (1e-12 if precision == FP64 else 2e-5)
