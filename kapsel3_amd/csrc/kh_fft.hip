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
#include "kh_zwave.h"

namespace kh {


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
// Block of 2P consecutive real rows = P complex lines (rows 2p and 2p+1 are
// the real and imaginary parts of line p).
template <typename T, int L>
struct ZPass {
    static constexpr int NT  = PassThreads<T>::value;
    static constexpr int P   = (FftTile<T>::ZELEM / L) > 0 ? FftTile<T>::ZELEM / L : 1;
    static constexpr int NZH = L / 2 + 1;
    static constexpr int BS  = ColStride<T, P>::value;
    static constexpr int SZ  = (L * BS > P * (L + 2)) ? L * BS : P * (L + 2);  // FFT tile or staged rows
    static constexpr int NR  = (2 * P * L + NT - 1) / NT;                     // real points per thread
    static constexpr int NC  = (P * L + NT - 1) / NT;                         // complex points per thread
};

template <typename T, int L>
__global__ __launch_bounds__(PassThreads<T>::value) void k_fft_z_r2c(const T *__restrict__ in,
                                                                     cpx<T> *__restrict__ out, long long nrows,
                                                                     int pitch) {
    using Z = ZPass<T, L>;
    constexpr int NT = Z::NT, P = Z::P, NZH = Z::NZH, BS = Z::BS;
    __shared__ cpx<T> s[Z::SZ], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const long long row0 = (long long)blockIdx.x * (2 * P);
    const T        *src  = in + row0 * L;
    T               v[Z::NR];
#pragma unroll
    for (int i = 0; i < Z::NR; i++) {
        const int e = threadIdx.x + i * NT;
        v[i]        = (e < 2 * P * L && row0 + e / L < nrows) ? src[e] : (T)0;
    }
#pragma unroll
    for (int i = 0; i < Z::NR; i++) {
        const int e = threadIdx.x + i * NT;
        if (e < 2 * P * L) {
            const int rl = e / L, n = e - rl * L;
            reinterpret_cast<T *>(&s[n * BS + (rl >> 1)])[rl & 1] = v[i];
        }
    }
    __syncthreads();
    fft_cols_ip<T, L, P, BS, NT, true>(s, tw);
    cpx<T> *dst = out + row0 * pitch;
    for (int e = threadIdx.x; e < 2 * P * NZH; e += NT) {
        const int rl = e / NZH, k = e - rl * NZH;
        if (row0 + rl >= nrows) break;
        const int    p  = rl >> 1;
        const cpx<T> Zk = s[k * BS + p];
        const cpx<T> Zm = cconj(s[((L - k) % L) * BS + p]);
        cpx<T>       o;
        if (rl & 1) {  // -i/2 (Z - Zm)
            const cpx<T> d = csub(Zk, Zm);
            o              = cmk(d.y * (T)0.5, -d.x * (T)0.5);
        } else {
            o = cscale(cadd(Zk, Zm), (T)0.5);
        }
        dst[rl * pitch + k] = o;
    }
}

template <typename T, int L>
__global__ __launch_bounds__(PassThreads<T>::value) void k_fft_z_c2r(ZIO<T> io, long long nrows, T scale,
                                                                     int kzvalid, int pitch, Grid bg) {
    if (batch_skip(bg)) return;
    const cpx<T> *__restrict__ in  = brun(io.in[blockIdx.y], bg);  // one component per grid row
    T *__restrict__            out = brun(io.out[blockIdx.y], bg);
    using Z = ZPass<T, L>;
    constexpr int NT = Z::NT, P = Z::P, NZH = Z::NZH, BS = Z::BS;
    __shared__ cpx<T> s[Z::SZ], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const long long row0 = (long long)blockIdx.x * (2 * P);
    const cpx<T>   *src  = in + row0 * pitch;
    for (int e = threadIdx.x; e < 2 * P * NZH; e += NT) {  // modes kz >= kzvalid are zero
        const int rl = e / NZH, k = e - rl * NZH;
        s[e]         = (row0 + rl < nrows && k < kzvalid) ? src[rl * pitch + k] : cmk((T)0, (T)0);
    }
    __syncthreads();
    cpx<T> zv[Z::NC];
#pragma unroll
    for (int i = 0; i < Z::NC; i++) {
        const int e = threadIdx.x + i * NT;
        if (e < P * L) zv[i] = hermitian_pair<T, L>(s, e / L, e % L);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < Z::NC; i++) {
        const int e = threadIdx.x + i * NT;
        if (e < P * L) s[(e % L) * BS + e / L] = zv[i];
    }
    __syncthreads();
    fft_cols_ip<T, L, P, BS, NT, false>(s, tw);
    T *dst = out + row0 * L;
    for (int e = threadIdx.x; e < 2 * P * L; e += NT) {
        const int rl = e / L, n = e - rl * L;
        if (row0 + rl >= nrows) break;
        const cpx<T> z = s[n * BS + (rl >> 1)];
        dst[e]         = ((rl & 1) ? z.y : z.x) * scale;
    }
}

// ------------------------------------------------------------ y / x pass
// Lines of L points at stride `lstride`; a tile is C consecutive columns.
// Tile t of outer block o starts at data + o*ostride + t*C; `ncols` columns
// are valid per outer block.
template <typename T, int L, bool FWD>
__global__ __launch_bounds__(PassThreads<T>::value) void k_fft_strided(cpx<T> *__restrict__ data, long long lstride,
                                                                       int ncols, long long ostride, int tiles) {
    constexpr int NT  = PassThreads<T>::value;
    constexpr int C   = FftTile<T>::C;
    constexpr int BS  = ColStride<T, C>::value;
    constexpr int NPT = (C * L + NT - 1) / NT;
    __shared__ cpx<T> s[L * BS], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const int o = blockIdx.x / tiles, t = blockIdx.x - o * tiles;
    const int c0 = t * C;
    cpx<T>   *base = data + (long long)o * ostride + c0;
    cpx<T>    v[NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int e = threadIdx.x + i * NT;
        const int n = e / C, c = e - n * C;
        v[i] = (e < C * L && c0 + c < ncols) ? base[(long long)n * lstride + c] : cmk((T)0, (T)0);
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int e = threadIdx.x + i * NT;
        if (e < C * L) s[(e / C) * BS + e % C] = v[i];
    }
    __syncthreads();
    fft_cols_ip<T, L, C, BS, NT, FWD>(s, tw);
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int e = threadIdx.x + i * NT;
        const int n = e / C, c = e - n * C;
        if (e < C * L && c0 + c < ncols) base[(long long)n * lstride + c] = s[n * BS + c];
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
static hipError_t launch_z_r2c(int L, const T *in, cpx<T> *out, long long nrows, int pitch, hipStream_t s) {
    switch (L) {
#define KH_CASE(n)                                                                       \
    case n: {                                                                            \
        constexpr int P  = ZPass<T, n>::P;                                               \
        long long     nb = (nrows + 2 * P - 1) / (2 * P);                                \
        hipLaunchKernelGGL((k_fft_z_r2c<T, n>), dim3((unsigned)nb), dim3(ZPass<T, n>::NT), 0, s, in, out, nrows, pitch); \
        return hipGetLastError();                                                        \
    }
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
static hipError_t launch_z_c2r(int L, const ZIO<T> &io, int ncomp, long long nrows, T scale, int kzvalid,
                               int pitch, hipStream_t s, const Grid &bg) {
    switch (L) {
#define KH_CASE(n)                                                                       \
    case n: {                                                                            \
        constexpr int P  = ZPass<T, n>::P;                                               \
        long long     nb = (nrows + 2 * P - 1) / (2 * P);                                \
        hipLaunchKernelGGL((k_fft_z_c2r<T, n>), dim3((unsigned)nb, ncomp, batch_z(bg)), dim3(ZPass<T, n>::NT), 0, s, io, nrows, scale, kzvalid, pitch, bg); \
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
        hipLaunchKernelGGL((k_fft_strided<T, n, FWD>), dim3((unsigned)nb), dim3(PassThreads<T>::value), 0, s, data, lstride, ncols, \
                           ostride, tiles);                                                                     \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t fft_z_r2c(const Grid &g, const T *in, cpx<T> *out, hipStream_t s, int kzvalid) {
    if (zw_supported(g.nz)) return zw_r2c<T>(g.nz, in, out, (long long)g.nx * g.ny, g.nzp, s, kzvalid);
    return launch_z_r2c<T>(g.nz, in, out, (long long)g.nx * g.ny, g.nzp, s);
}
template <typename T>
hipError_t fft_z_c2r(const Grid &g, const cpx<T> *in, T *out, hipStream_t s, int kzvalid) {
    ZIO<T> io{};
    io.in[0]  = in;
    io.out[0] = out;
    return fft_z_c2r_multi<T>(g, io, 1, s, kzvalid);
}
template <typename T>
hipError_t fft_z_c2r_multi(const Grid &g, const ZIO<T> &io, int ncomp, hipStream_t s, int kzvalid, const T *mask,
                           Occ occ) {
    const T   scale = (T)(1.0 / ((double)g.nx * (double)g.ny * (double)g.nz));
    const int kzv   = kzvalid < 0 ? g.nzh : kzvalid;
    if (zw_supported(g.nz))
        return zw_c2r<T>(g.nz, io, ncomp, (long long)g.nx * g.ny, scale, kzv, g.nzp, s, mask, nullptr, &g, occ);
    if (occ.w) return hipErrorInvalidValue;
    return launch_z_c2r<T>(g.nz, io, ncomp, (long long)g.nx * g.ny, scale, kzv, g.nzp, s, g);
}
// y: lines along y at stride nzp, columns kz < nzh, one outer block per x
template <typename T, bool FWD>
hipError_t fft_y(const Grid &g, cpx<T> *data, hipStream_t s) {
    return launch_strided<T, FWD>(g.ny, data, g.nzp, g.nzh, (long long)g.ny * g.nzp, g.nx, s);
}
// x: lines along x at stride ny*nzp, columns (y,kz) flattened (the row
// padding columns are transformed too: they hold zeros)
template <typename T, bool FWD>
hipError_t fft_x(const Grid &g, cpx<T> *data, hipStream_t s) {
    const int M = g.ny * g.nzp;
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
    template hipError_t fft_z_r2c<T>(const Grid &, const T *, cpx<T> *, hipStream_t, int); \
    template hipError_t fft_z_c2r<T>(const Grid &, const cpx<T> *, T *, hipStream_t, int); \
    template hipError_t fft_z_c2r_multi<T>(const Grid &, const ZIO<T> &, int, hipStream_t, int, const T *, Occ); \
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
