// kh_fft.hip -- 3D real FFT passes on HBM-resident fields (gfx950).
//
// Replaces Ooura's rdft3d/rdft3dsort behind A2a_k / A_k2a
// (fft_wrapper.h:641-681, fftsg3d.c:554-709). Layouts on the device:
//   real     R[x][y][z]   z fastest, NX*NY*NZ values (no padding)
//   spectral S[x][y][kz]  complex, kz = 0..NZ/2 fastest, NX*NY*(NZ/2+1)
// The spectral layout is the reference's ([i][j][2kz] as (re,im)), so host
// transfers of spectral fields are plain copies.
//
// Forward = z-pass (two real rows packed into one complex row, r2c), then
// in-place complex passes along y and along x. Inverse = x, y, then z c2r,
// with the 1/(NX NY NZ) normalisation folded into the last pass. Every pass
// reads and writes each field element once, staging a tile in LDS:
//   z pass : a contiguous block of rows (fully coalesced both ways)
//   y/x    : a tile of C neighbouring columns (C*elem = 128 B per row)
#include "kh_fft.h"
#include "kh_fft_engine.h"

namespace kh {

constexpr int FFT_NT = 256;

template <typename T>
struct FftTile;
template <>
struct FftTile<double> {
    static constexpr int C     = 8;     // columns per y/x tile (128 B rows)
    static constexpr int ZELEM = 2048;  // complex points per z-pass block
};
template <>
struct FftTile<float> {
    static constexpr int C     = 16;
    static constexpr int ZELEM = 4096;
};

// ---------------------------------------------------------------- z pass
template <typename T, int L>
__global__ __launch_bounds__(FFT_NT) void k_fft_z_r2c(const T *__restrict__ in, cpx<T> *__restrict__ out,
                                                      long long nrows) {
    constexpr int P   = (FftTile<T>::ZELEM / L) > 0 ? FftTile<T>::ZELEM / L : 1;  // complex lines per WG
    constexpr int NZH = L / 2 + 1;
    __shared__ cpx<T> s0[P * L], s1[P * L], tw[L];
    fill_twiddles<T, L, FFT_NT>(tw);
    const long long row0 = (long long)blockIdx.x * (2 * P);
    const T        *src  = in + row0 * L;
    for (int e = threadIdx.x; e < 2 * P * L; e += FFT_NT) {
        const int rl = e / L, n = e - rl * L;
        T         v  = (row0 + rl < nrows) ? src[e] : (T)0;
        T        *d  = reinterpret_cast<T *>(&s0[(rl >> 1) * L + n]);
        d[rl & 1]    = v;
    }
    __syncthreads();
    cpx<T> *res = fft_lines<T, L, P, L, FFT_NT, true>(s0, s1, tw);
    cpx<T> *dst = out + row0 * NZH;
    for (int e = threadIdx.x; e < 2 * P * NZH; e += FFT_NT) {
        const int rl = e / NZH, k = e - rl * NZH;
        if (row0 + rl >= nrows) break;
        const int    p  = rl >> 1;
        const cpx<T> Z  = res[p * L + k];
        const cpx<T> Zm = cconj(res[p * L + ((L - k) % L)]);
        cpx<T>       v;
        if (rl & 1) {  // -i/2 (Z - Zm)
            cpx<T> d = csub(Z, Zm);
            v        = cmk(d.y * (T)0.5, -d.x * (T)0.5);
        } else {
            v = cscale(cadd(Z, Zm), (T)0.5);
        }
        dst[e] = v;
    }
}

template <typename T, int L>
__global__ __launch_bounds__(FFT_NT) void k_fft_z_c2r(const cpx<T> *__restrict__ in, T *__restrict__ out,
                                                      long long nrows, T scale, int kzvalid) {
    constexpr int P   = (FftTile<T>::ZELEM / L) > 0 ? FftTile<T>::ZELEM / L : 1;
    constexpr int NZH = L / 2 + 1;
    __shared__ cpx<T> s0[P * L], s1[P * (L + 2)], tw[L];
    fill_twiddles<T, L, FFT_NT>(tw);
    const long long row0 = (long long)blockIdx.x * (2 * P);
    const cpx<T>   *src  = in + row0 * NZH;
    for (int e = threadIdx.x; e < 2 * P * NZH; e += FFT_NT) {  // modes kz >= kzvalid are zero
        const int rl = e / NZH, k = e - rl * NZH;
        s1[e]        = (row0 + rl < nrows && k < kzvalid) ? src[e] : cmk((T)0, (T)0);
    }
    __syncthreads();
    // two Hermitian rows A (even), B (odd) -> Z = A + iB over the full line;
    // imaginary parts of kz = 0 and kz = L/2 are dropped, as any c2r does.
    for (int e = threadIdx.x; e < P * L; e += FFT_NT) {
        const int p = e / L, n = e - p * L;
        const int k = (n <= L / 2) ? n : L - n;
        cpx<T>    A = s1[(2 * p) * NZH + k], B = s1[(2 * p + 1) * NZH + k];
        if (k == 0 || k == L / 2) {
            A.y = 0;
            B.y = 0;
        }
        if (n > L / 2) {
            A = cconj(A);
            B = cconj(B);
        }
        s0[e] = cmk(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    cpx<T> *res = fft_lines<T, L, P, L, FFT_NT, false>(s0, s1, tw);
    T      *dst = out + row0 * L;
    for (int e = threadIdx.x; e < 2 * P * L; e += FFT_NT) {
        const int rl = e / L, n = e - rl * L;
        if (row0 + rl >= nrows) break;
        const cpx<T> z = res[(rl >> 1) * L + n];
        dst[e]         = ((rl & 1) ? z.y : z.x) * scale;
    }
}

// ------------------------------------------------------------ y / x pass
// Lines of L points at stride `lstride`; a tile is C consecutive columns.
// Tile t of outer block o starts at data + o*ostride + t*C; `ncols` columns
// are valid per outer block.
template <typename T, int L, bool FWD>
__global__ __launch_bounds__(FFT_NT) void k_fft_strided(cpx<T> *__restrict__ data, long long lstride, int ncols,
                                                        long long ostride, int tiles) {
    constexpr int C  = FftTile<T>::C;
    constexpr int LS = L + 1;
    __shared__ cpx<T> s0[C * LS], s1[C * LS], tw[L];
    fill_twiddles<T, L, FFT_NT>(tw);
    const int o = blockIdx.x / tiles, t = blockIdx.x - o * tiles;
    const int c0 = t * C;
    cpx<T>   *base = data + (long long)o * ostride + c0;
    for (int e = threadIdx.x; e < C * L; e += FFT_NT) {
        const int n = e / C, c = e - n * C;
        s0[c * LS + n] = (c0 + c < ncols) ? base[(long long)n * lstride + c] : cmk((T)0, (T)0);
    }
    __syncthreads();
    cpx<T> *res = fft_lines<T, L, C, LS, FFT_NT, FWD>(s0, s1, tw);
    for (int e = threadIdx.x; e < C * L; e += FFT_NT) {
        const int n = e / C, c = e - n * C;
        if (c0 + c < ncols) base[(long long)n * lstride + c] = res[c * LS + n];
    }
}

// ------------------------------------------------------------ launchers
#define KH_FOR_EACH_L(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

bool fft_length_supported(int L) {
    switch (L) {
#define KH_CASE(n) case n:
        KH_FOR_EACH_L(KH_CASE)
        return true;
#undef KH_CASE
        default: return false;
    }
}

template <typename T>
static hipError_t launch_z_r2c(int L, const T *in, cpx<T> *out, long long nrows, hipStream_t s) {
    switch (L) {
#define KH_CASE(n)                                                                       \
    case n: {                                                                            \
        constexpr int P  = (FftTile<T>::ZELEM / n) > 0 ? FftTile<T>::ZELEM / n : 1;      \
        long long     nb = (nrows + 2 * P - 1) / (2 * P);                                \
        hipLaunchKernelGGL((k_fft_z_r2c<T, n>), dim3((unsigned)nb), dim3(FFT_NT), 0, s, in, out, nrows); \
        return hipGetLastError();                                                        \
    }
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
static hipError_t launch_z_c2r(int L, const cpx<T> *in, T *out, long long nrows, T scale, int kzvalid,
                               hipStream_t s) {
    switch (L) {
#define KH_CASE(n)                                                                       \
    case n: {                                                                            \
        constexpr int P  = (FftTile<T>::ZELEM / n) > 0 ? FftTile<T>::ZELEM / n : 1;      \
        long long     nb = (nrows + 2 * P - 1) / (2 * P);                                \
        hipLaunchKernelGGL((k_fft_z_c2r<T, n>), dim3((unsigned)nb), dim3(FFT_NT), 0, s, in, out, nrows, scale, kzvalid); \
        return hipGetLastError();                                                        \
    }
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T, bool FWD>
static hipError_t launch_strided(int L, cpx<T> *data, long long lstride, int ncols, long long ostride, int nouter,
                                 hipStream_t s) {
    constexpr int C     = FftTile<T>::C;
    const int     tiles = (ncols + C - 1) / C;
    const long long nb  = (long long)tiles * nouter;
    switch (L) {
#define KH_CASE(n)                                                                                              \
    case n:                                                                                                     \
        hipLaunchKernelGGL((k_fft_strided<T, n, FWD>), dim3((unsigned)nb), dim3(FFT_NT), 0, s, data, lstride, ncols, \
                           ostride, tiles);                                                                     \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t fft_z_r2c(const Grid &g, const T *in, cpx<T> *out, hipStream_t s) {
    return launch_z_r2c<T>(g.nz, in, out, (long long)g.nx * g.ny, s);
}
template <typename T>
hipError_t fft_z_c2r(const Grid &g, const cpx<T> *in, T *out, hipStream_t s, int kzvalid) {
    const T scale = (T)(1.0 / ((double)g.nx * (double)g.ny * (double)g.nz));
    return launch_z_c2r<T>(g.nz, in, out, (long long)g.nx * g.ny, scale, kzvalid < 0 ? g.nzh : kzvalid, s);
}
// y: lines along y at stride nzh, columns kz, one outer block per x
template <typename T, bool FWD>
hipError_t fft_y(const Grid &g, cpx<T> *data, hipStream_t s) {
    return launch_strided<T, FWD>(g.ny, data, g.nzh, g.nzh, (long long)g.ny * g.nzh, g.nx, s);
}
// x: lines along x at stride ny*nzh, columns (y,kz) flattened
template <typename T, bool FWD>
hipError_t fft_x(const Grid &g, cpx<T> *data, hipStream_t s) {
    const int M = g.ny * g.nzh;
    return launch_strided<T, FWD>(g.nx, data, M, M, 0, 1, s);
}

template <typename T>
hipError_t fft3d_r2c(const Grid &g, const T *real_in, cpx<T> *spec_out, hipStream_t s) {
    hipError_t e;
    if ((e = fft_z_r2c<T>(g, real_in, spec_out, s)) != hipSuccess) return e;
    if ((e = fft_y<T, true>(g, spec_out, s)) != hipSuccess) return e;
    return fft_x<T, true>(g, spec_out, s);
}

template <typename T>
hipError_t fft3d_c2r(const Grid &g, cpx<T> *spec_inout, T *real_out, hipStream_t s) {
    hipError_t e;
    if ((e = fft_x<T, false>(g, spec_inout, s)) != hipSuccess) return e;
    if ((e = fft_y<T, false>(g, spec_inout, s)) != hipSuccess) return e;
    return fft_z_c2r<T>(g, spec_inout, real_out, s);
}

#define KH_INST(T)                                                                   \
    template hipError_t fft_z_r2c<T>(const Grid &, const T *, cpx<T> *, hipStream_t); \
    template hipError_t fft_z_c2r<T>(const Grid &, const cpx<T> *, T *, hipStream_t, int); \
    template hipError_t fft_y<T, true>(const Grid &, cpx<T> *, hipStream_t);          \
    template hipError_t fft_y<T, false>(const Grid &, cpx<T> *, hipStream_t);         \
    template hipError_t fft_x<T, true>(const Grid &, cpx<T> *, hipStream_t);          \
    template hipError_t fft_x<T, false>(const Grid &, cpx<T> *, hipStream_t);
KH_INST(double)
KH_INST(float)
#undef KH_INST

template hipError_t fft3d_r2c<double>(const Grid &, const double *, cpx<double> *, hipStream_t);
template hipError_t fft3d_r2c<float>(const Grid &, const float *, cpx<float> *, hipStream_t);
template hipError_t fft3d_c2r<double>(const Grid &, cpx<double> *, double *, hipStream_t);
template hipError_t fft3d_c2r<float>(const Grid &, cpx<float> *, float *, hipStream_t);

}  // namespace kh
