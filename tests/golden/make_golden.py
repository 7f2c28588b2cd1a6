"""Writes the golden fixtures in this directory from the reference's own
code, compiled under oracle/_ref from fftsg.c, fftsg3d.c and mt19937ar.c
where they lie in the reference tree (`make -C oracle ref`).

  ooura_fft.npz  seeded real fields and their A2a_k / A_k2a transforms by
                 the reference's Ooura rdft3d with the call sequence of
                 fft_wrapper.h:641-681 (default, non-FFTW branch)
  mt19937.npz    genrand_int32 / genrand_real3 after init_genrand(s)

Run in the build container only (it needs the reference tree):
    python tests/golden/make_golden.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "libooura_ref.so")


def main():
    L = C.CDLL(REF)
    L.ref_A2a_k.argtypes = [C.c_int] * 3 + [C.c_void_p]
    L.ref_A_k2a.argtypes = [C.c_int] * 3 + [C.c_void_p]
    L.ref_mt_seed.argtypes = [C.c_ulong]
    L.ref_mt_real3.restype = C.c_double
    L.ref_mt_int32.restype = C.c_ulong
    out = {}
    for (nx, ny, nz) in [(8, 8, 8), (16, 8, 32), (16, 16, 16)]:
        nz_ = 2 * (nz // 2 + 1)
        rng = np.random.default_rng(nx * 10000 + ny * 100 + nz)
        a = np.zeros((nx, ny, nz_))
        a[:, :, :nz] = rng.standard_normal((nx, ny, nz))
        f = a.copy()
        L.ref_A2a_k(nx, ny, nz, f.ctypes.data)
        b = f.copy()
        L.ref_A_k2a(nx, ny, nz, b.ctypes.data)
        tag = f"{nx}x{ny}x{nz}"
        out[f"in_{tag}"] = a
        out[f"fwd_{tag}"] = f
        out[f"inv_{tag}"] = b[:, :, :nz]
    np.savez_compressed(os.path.join(HERE, "ooura_fft.npz"), **out)
    mt = {}
    for s in (0, 5489, 10):
        L.ref_mt_seed(s)
        mt[f"int32_seed{s}"] = np.array([L.ref_mt_int32() for _ in range(64)], dtype=np.uint64)
        L.ref_mt_seed(s)
        mt[f"real3_seed{s}"] = np.array([L.ref_mt_real3() for _ in range(64)])
    np.savez_compressed(os.path.join(HERE, "mt19937.npz"), **mt)


if __name__ == "__main__":
    main()
