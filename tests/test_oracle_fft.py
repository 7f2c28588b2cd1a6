"""Pins the oracle's FFT (oracle/kref_fft.c) to the reference's own Ooura
transform, to numpy, and to closed forms. CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import kref

HERE = os.path.dirname(os.path.abspath(__file__))
REF_LIB = os.path.join(HERE, "..", "oracle", "_ref", "libooura_ref.so")


def _field(nx, ny, nz, seed):
    a = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    a[:, :, :nz] = np.random.default_rng(seed).standard_normal((nx, ny, nz))
    return a


def test_golden_ooura_vectors():
    """A2a_k / A_k2a of the reference (Ooura rdft3d, fft_wrapper.h:641-681)
    on committed fixtures."""
    d = np.load(os.path.join(HERE, "golden", "ooura_fft.npz"))
    tags = sorted(k[3:] for k in d.files if k.startswith("in_"))
    assert tags
    for tag in tags:
        a, fwd, inv = d["in_" + tag], d["fwd_" + tag], d["inv_" + tag]
        nz = a.shape[2] - 2
        o = kref.fft_forward(a)
        assert np.abs(o - fwd).max() / np.abs(fwd).max() < 1e-14, tag
        b = kref.fft_inverse(fwd)
        assert np.abs(b[:, :, :nz] - inv).max() < 1e-14, tag


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference Ooura build (oracle/_ref) not present")
@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 32, 16), (8, 64, 32)])
def test_against_reference_build(shape):
    L = C.CDLL(REF_LIB)
    L.ref_A2a_k.argtypes = [C.c_int] * 3 + [C.c_void_p]
    L.ref_A_k2a.argtypes = [C.c_int] * 3 + [C.c_void_p]
    nx, ny, nz = shape
    a = _field(nx, ny, nz, 3)
    f = a.copy()
    L.ref_A2a_k(nx, ny, nz, f.ctypes.data)
    o = kref.fft_forward(a)
    assert np.abs(o - f).max() / np.abs(f).max() < 1e-14
    b = f.copy()
    L.ref_A_k2a(nx, ny, nz, b.ctypes.data)
    ob = kref.fft_inverse(o)
    assert np.abs(ob[:, :, :nz] - b[:, :, :nz]).max() < 1e-13


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 24, 48), (48, 48, 48), (96, 32, 16), (16, 16, 96), (64, 64, 64)])
def test_against_numpy(shape):
    nx, ny, nz = shape
    a = _field(nx, ny, nz, 5)
    o = kref.fft_forward(a)
    ref = np.fft.rfftn(a[:, :, :nz])
    oc = o[:, :, 0::2] + 1j * o[:, :, 1::2]
    assert np.abs(oc - ref).max() / np.abs(ref).max() < 1e-14
    back = kref.fft_inverse(o)
    assert np.abs(back[:, :, :nz] - a[:, :, :nz]).max() < 1e-13


def test_single_mode_closed_form():
    """cos(2 pi (m.x)/N) transforms to N^3/2 at +-m and 0 elsewhere; the
    padding columns are ignored on input."""
    n = 16
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    m = (3, -5, 2)
    a = np.full((n, n, n + 2), 7.0)  # garbage in padding
    a[:, :, :n] = np.cos(2 * np.pi * (m[0] * X + m[1] * Y + m[2] * Z) / n)
    o = kref.fft_forward(a)
    oc = o[:, :, 0::2] + 1j * o[:, :, 1::2]
    expect = np.zeros_like(oc)
    expect[m[0] % n, m[1] % n, m[2]] = n ** 3 / 2
    assert np.abs(oc - expect).max() < 1e-10
    # sine: -i N^3/2 at +m (forward kernel exp(-i k.x))
    a[:, :, :n] = np.sin(2 * np.pi * (m[0] * X + m[1] * Y + m[2] * Z) / n)
    oc = kref.fft_forward(a)
    oc = oc[:, :, 0::2] + 1j * oc[:, :, 1::2]
    assert abs(oc[m[0] % n, m[1] % n, m[2]] - (-0.5j * n ** 3)) < 1e-10


def test_parseval_and_linearity():
    n = 24
    a, b = _field(n, n, n, 1), _field(n, n, n, 2)
    fa, fb, fab = kref.fft_forward(a), kref.fft_forward(b), kref.fft_forward(2.5 * a - b)
    assert np.abs(fab - (2.5 * fa - fb)).max() < 1e-11
    c = fa[:, :, 0::2] + 1j * fa[:, :, 1::2]
    w = np.ones(n // 2 + 1)
    w[1:n // 2] = 2.0  # Hermitian half-spectrum weights
    e_spec = (np.abs(c) ** 2 * w).sum() / n ** 3
    assert abs(e_spec - (a[:, :, :n] ** 2).sum()) / e_spec < 1e-13
