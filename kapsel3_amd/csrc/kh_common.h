// kh_common.h -- device-side basics shared by the SPM kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace kh {

// Complex value in registers/LDS/HBM: (re, im), 8- or 16-byte aligned pair.
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmk(T a, T b) {
    cpx<T> r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
    return cmk(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
    return cmk(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return cmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
    return cmk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) {
    return cmk(a.x, -a.y);
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) {
    return cmk(a.x * s, a.y * s);
}
// multiply by -i (FWD) or +i (inverse)
template <typename T, bool FWD>
__device__ __forceinline__ cpx<T> cmul_mi(cpx<T> a) {
    return FWD ? cmk(a.y, -a.x) : cmk(-a.y, a.x);
}

// Grid geometry of one run, passed by value to kernels.
struct Grid {
    int    nx, ny, nz;       // real-space points
    int    nzh;              // nz/2 + 1 complex modes along z
    int    nzp;              // row pitch of spectral arrays (complex elements): nzh rounded up to 128 bytes
    int    hnx, hny;         // nx/2, ny/2 (Calc_KX_Ooura: i > HNX -> i - NX)
    int    trn_x, trn_y, trn_z;  // (N+2)/3: kept |kx| < trn_x, |ky| < trn_y, kz < trn_z
    double wave_x, wave_y, wave_z;
    long long nreal;         // nx*ny*nz
    long long nspec;         // nx*ny*nzp (allocated spectral elements, row padding included)
    // A batch of runs of this grid stepped by the same launches (kh_batch):
    // grid z of a launch is the run, run b's buffers sit b * bstride bytes
    // after run 0's (equal device arenas), and a run with bactive[b] == 0 is
    // skipped. nbatch <= 1: one run.
    int        nbatch;
    long long  bstride;
    const int *bactive;
};

// Launch grid z of a pass over g's batch.
inline unsigned batch_z(const Grid &g) { return g.nbatch > 1 ? (unsigned)g.nbatch : 1u; }
// This block's run of the batch is not stepped.
__device__ __forceinline__ bool batch_skip(const Grid &g) { return g.bactive && !g.bactive[blockIdx.z]; }
// p, a buffer of run 0, moved to this block's run of the batch. The move
// is byte arithmetic on the pointer itself (never through an integer), so
// the compiler still sees a global-memory pointer and emits global_load /
// global_store rather than flat accesses, which would also count against
// the LDS wait counter.
template <typename P>
__device__ __forceinline__ P *brun(P *p, const Grid &g) {
    using C = typename std::conditional<std::is_const<P>::value, const char, char>::type;
    return p ? reinterpret_cast<P *>(reinterpret_cast<C *>(p) + (long long)blockIdx.z * g.bstride) : p;
}

// Occupancy of a run's 4-point vectors along z: bit (v & 63) of word
// row * words + (v >> 6) is set when some particle's phi may be nonzero in
// vector v (points 4v..4v+3) of real-space row `row` (the line along z at
// (x, y), row = x * ny + y). The phi scatter sets the bits; the last pass
// that reads them clears them again.
struct Occ {
    unsigned long long *w     = nullptr;
    int                 words = 0;
};
// the bits are kept for FP64 fields only unless this is set (in FP32 they
// cost the scatter as much as they save the masked passes)
#ifndef KH_OCC_FP32
#define KH_OCC_FP32 0
#endif
template <typename T>
constexpr bool occ_kept() { return sizeof(T) == 8 || KH_OCC_FP32; }
__device__ __forceinline__ bool occ_bit(const unsigned long long *w, int words, long long row, int v) {
    return (w[row * words + (v >> 6)] >> (v & 63)) & 1ull;
}
// set the bits of vectors a..b (a <= b) of one row
__device__ __forceinline__ void occ_mark(const Occ &o, int row, int a, int b) {
    for (int w = a >> 6; w <= (b >> 6); w++) {
        const int                lo = (a > 64 * w ? a : 64 * w) - 64 * w, hi = (b < 64 * w + 63 ? b : 64 * w + 63) - 64 * w;
        const unsigned long long m  = (hi == 63 ? ~0ull : (1ull << (hi + 1)) - 1ull) & ~((1ull << lo) - 1ull);
        atomicOr(o.w + (long long)row * o.words + w, m);
    }
}

__device__ __forceinline__ int kx_of(const Grid &g, int i) { return (i > g.hnx) ? i - g.nx : i; }
__device__ __forceinline__ int ky_of(const Grid &g, int j) { return (j > g.hny) ? j - g.ny : j; }
__device__ __forceinline__ bool kept1(int k, int trn) { return (k < 0 ? -k : k) < trn; }
__device__ __forceinline__ int  nkept(int trn) { return 2 * trn - 1; }
// o-th kept index of an axis of n points: 0..trn-1, then n-trn+1..n-1
__device__ __forceinline__ int kept_index(int o, int n, int trn) { return o < trn ? o : n - nkept(trn) + o; }
__device__ __forceinline__ bool kept_mode(const Grid &g, int kx, int ky, int kz) {
    return kz < g.trn_z && (ky < 0 ? -ky : ky) < g.trn_y && (kx < 0 ? -kx : kx) < g.trn_x;
}

}  // namespace kh
