// kh_fft_engine.h -- in-LDS batched 1D FFT for gfx950 (Stockham autosort).
//
// A workgroup holds B lines of L complex points in LDS (line-major, line
// stride LS) and transforms all of them together. Each stage of radix R is
// one sweep: every thread takes butterflies (line, j), loads R points at
// stride L/R, applies the stage twiddles, does the R-point DFT in registers
// and writes R outputs to the other LDS buffer in Stockham (self-sorting)
// order, so the result comes out in natural order with no bit reversal.
// Radices 8, 4, 2 and 3 cover every L = 2^a 3^b; the plan is fixed at
// compile time so all index arithmetic folds to shifts and multiplies.
#pragma once

#include "kh_common.h"

namespace kh {

struct RadixPlan {
    int n;       // number of stages
    int r[16];   // radix of stage s
    int ns[16];  // product of the radices before stage s
};

__host__ __device__ constexpr RadixPlan make_radix_plan(int L) {
    RadixPlan p{};
    int       e2 = 0, e3 = 0, m = L;
    while (m % 2 == 0) {
        m /= 2;
        e2++;
    }
    while (m % 3 == 0) {
        m /= 3;
        e3++;
    }
    int n8 = e2 / 3, n4 = 0, n2 = 0;
    int rem = e2 % 3;
    if (rem == 2) n4 = 1;
    if (rem == 1) {
        if (n8 > 0) {  // 8*2 -> 4*4
            n8--;
            n4 = 2;
        } else {
            n2 = 1;
        }
    }
    int k = 0;
    for (int i = 0; i < n8; i++) p.r[k++] = 8;
    for (int i = 0; i < n4; i++) p.r[k++] = 4;
    for (int i = 0; i < n2; i++) p.r[k++] = 2;
    for (int i = 0; i < e3; i++) p.r[k++] = 3;
    p.n     = k;
    int acc = 1;
    for (int i = 0; i < k; i++) {
        p.ns[i] = acc;
        acc *= p.r[i];
    }
    return p;
}

__host__ __device__ constexpr bool fft_size_ok(int L) {
    if (L < 2) return false;
    int m = L;
    while (m % 2 == 0) m /= 2;
    while (m % 3 == 0) m /= 3;
    return m == 1;
}

// ---------------------------------------------------------------- DFTs
template <typename T, int R, bool FWD>
struct Dft;

template <typename T, bool FWD>
struct Dft<T, 2, FWD> {
    __device__ __forceinline__ static void run(cpx<T> *v) {
        cpx<T> a = v[0], b = v[1];
        v[0]     = cadd(a, b);
        v[1]     = csub(a, b);
    }
};

template <typename T, bool FWD>
struct Dft<T, 3, FWD> {
    __device__ __forceinline__ static void run(cpx<T> *v) {
        const T s3 = (T)0.86602540378443864676372317075294;  // sqrt(3)/2
        cpx<T>  t1 = cadd(v[1], v[2]);
        cpx<T>  t2 = csub(v[0], cscale(t1, (T)0.5));
        cpx<T>  d  = cscale(csub(v[1], v[2]), s3);
        cpx<T>  md = cmul_mi<T, FWD>(d);  // -i*s3*(v1-v2) forward
        v[0]       = cadd(v[0], t1);
        v[1]       = cadd(t2, md);
        v[2]       = csub(t2, md);
    }
};

template <typename T, bool FWD>
struct Dft<T, 4, FWD> {
    __device__ __forceinline__ static void run(cpx<T> *v) {
        cpx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        cpx<T> c = cadd(v[1], v[3]), d = cmul_mi<T, FWD>(csub(v[1], v[3]));
        v[0]     = cadd(a, c);
        v[1]     = cadd(b, d);
        v[2]     = csub(a, c);
        v[3]     = csub(b, d);
    }
};

template <typename T, bool FWD>
struct Dft<T, 8, FWD> {
    __device__ __forceinline__ static void run(cpx<T> *v) {
        const T h = (T)0.70710678118654752440084436210485;  // 1/sqrt(2)
        cpx<T>  e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T>  o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4, FWD>::run(e);
        Dft<T, 4, FWD>::run(o);
        // o[k] *= W8^k, W8 = exp(-+ 2 pi i / 8)
        cpx<T> o1 = FWD ? cmk((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h)
                        : cmk((o[1].x - o[1].y) * h, (o[1].y + o[1].x) * h);
        cpx<T> o2 = cmul_mi<T, FWD>(o[2]);
        cpx<T> o3 = FWD ? cmk((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h)
                        : cmk(-(o[3].x + o[3].y) * h, (o[3].x - o[3].y) * h);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// tw[n] = exp(-2 pi i n / L), n < L, filled by the whole workgroup.
template <typename T, int L, int NT>
__device__ __forceinline__ void fill_twiddles(cpx<T> *tw) {
    for (int n = threadIdx.x; n < L; n += NT) {
        double s, c;
        sincospi(2.0 * (double)n / (double)L, &s, &c);
        tw[n] = cmk((T)c, (T)(-s));
    }
}

template <typename T, int L, int R, int Ns, int B, int LS, int NT, bool FWD>
__device__ __forceinline__ void fft_stage(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out,
                                          const cpx<T> *__restrict__ tw) {
    constexpr int m       = L / R;
    constexpr int nbf     = B * m;
    constexpr int tstride = L / (Ns * R);
    for (int id = threadIdx.x; id < nbf; id += NT) {
        const int     line = id / m;
        const int     j    = id - line * m;
        const int     k    = j % Ns;
        cpx<T>        v[R];
        const cpx<T> *src = in + line * LS + j;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[r * m];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const cpx<T> w = tw[r * k * tstride];
                v[r]           = FWD ? cmul(v[r], w) : cmulc(v[r], w);
            }
        }
        Dft<T, R, FWD>::run(v);
        cpx<T> *dst = out + line * LS + (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dst[r * Ns] = v[r];
    }
}

template <typename T, int L, int B, int LS, int NT, bool FWD, int S>
__device__ __forceinline__ cpx<T> *fft_stages(cpx<T> *a, cpx<T> *b, const cpx<T> *tw) {
    constexpr RadixPlan P = make_radix_plan(L);
    if constexpr (S < P.n) {
        fft_stage<T, L, P.r[S], P.ns[S], B, LS, NT, FWD>(a, b, tw);
        __syncthreads();
        return fft_stages<T, L, B, LS, NT, FWD, S + 1>(b, a, tw);
    } else {
        return a;
    }
}

// Transforms B lines held in `a` (LDS, line stride LS); `b` is scratch of
// the same size. Returns the buffer holding the result. Caller must
// __syncthreads() before the call (input complete).
template <typename T, int L, int B, int LS, int NT, bool FWD>
__device__ __forceinline__ cpx<T> *fft_lines(cpx<T> *a, cpx<T> *b, const cpx<T> *tw) {
    static_assert(fft_size_ok(L), "FFT length must be 2^a 3^b");
    return fft_stages<T, L, B, LS, NT, FWD, 0>(a, b, tw);
}

// ------------------------------------------------------------------------
// Line-fastest variant: point n of line c lives at LDS index n*BS + c
// (BS >= B, padded). Butterfly id -> (line = id % B, j = id / B), so the
// lanes of a wavefront work on neighbouring lines at the same butterfly:
// their reads and writes fall on consecutive LDS words in every stage, and
// the row padding staggers the rows a radix-R stage writes R rows apart.
template <typename T, int L, int R, int Ns, int B, int BS, int NT, bool FWD>
__device__ __forceinline__ void fft_stage_cf(const cpx<T> *__restrict__ in, cpx<T> *__restrict__ out,
                                             const cpx<T> *__restrict__ tw) {
    constexpr int m       = L / R;
    constexpr int nbf     = B * m;
    constexpr int tstride = L / (Ns * R);
    for (int id = threadIdx.x; id < nbf; id += NT) {
        const int     j    = id / B;
        const int     line = id - j * B;
        const int     k    = j % Ns;
        cpx<T>        v[R];
        const cpx<T> *src  = in + j * BS + line;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[r * m * BS];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const cpx<T> w = tw[r * k * tstride];
                v[r]           = FWD ? cmul(v[r], w) : cmulc(v[r], w);
            }
        }
        Dft<T, R, FWD>::run(v);
        cpx<T> *dst = out + ((j / Ns) * Ns * R + k) * BS + line;
#pragma unroll
        for (int r = 0; r < R; r++) dst[r * Ns * BS] = v[r];
    }
}

template <typename T, int L, int B, int BS, int NT, bool FWD, int S>
__device__ __forceinline__ cpx<T> *fft_stages_cf(cpx<T> *a, cpx<T> *b, const cpx<T> *tw) {
    constexpr RadixPlan P = make_radix_plan(L);
    if constexpr (S < P.n) {
        fft_stage_cf<T, L, P.r[S], P.ns[S], B, BS, NT, FWD>(a, b, tw);
        __syncthreads();
        return fft_stages_cf<T, L, B, BS, NT, FWD, S + 1>(b, a, tw);
    } else {
        return a;
    }
}

// B lines, point n of line c at a[n*BS + c]; b is scratch of the same size.
// Returns the buffer holding the result; caller syncs before the call.
template <typename T, int L, int B, int BS, int NT, bool FWD>
__device__ __forceinline__ cpx<T> *fft_cols(cpx<T> *a, cpx<T> *b, const cpx<T> *tw) {
    static_assert(fft_size_ok(L), "FFT length must be 2^a 3^b");
    return fft_stages_cf<T, L, B, BS, NT, FWD, 0>(a, b, tw);
}

// Row stride for B lines of element type T: one extra element staggers the
// rows a stage writes R apart across the LDS banks.
template <typename T, int B>
struct ColStride {
    static constexpr int value = B + 1;
};

// In-place, register-staged variant of fft_cols: each stage reads every
// butterfly's R points into registers, synchronises, then writes the
// outputs back into the same buffer. One LDS buffer instead of two doubles
// the workgroups a CU can hold, which is what hides HBM latency here.
template <typename T, int L, int R, int Ns, int B, int BS, int NT, bool FWD>
__device__ __forceinline__ void fft_stage_ip(cpx<T> *__restrict__ buf, const cpx<T> *__restrict__ tw) {
    constexpr int m       = L / R;
    constexpr int nbf     = B * m;
    constexpr int NB      = (nbf + NT - 1) / NT;
    constexpr int tstride = L / (Ns * R);
    cpx<T>        v[NB][R];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int id = threadIdx.x + t * NT;
        if (NB * NT == nbf || id < nbf) {
            const int     j    = id / B;
            const int     line = id - j * B;
            const cpx<T> *src  = buf + j * BS + line;
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = src[r * m * BS];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NB; t++) {
        const int id = threadIdx.x + t * NT;
        if (NB * NT == nbf || id < nbf) {
            const int j    = id / B;
            const int line = id - j * B;
            const int k    = j % Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const cpx<T> w = tw[r * k * tstride];
                    v[t][r]        = FWD ? cmul(v[t][r], w) : cmulc(v[t][r], w);
                }
            }
            Dft<T, R, FWD>::run(v[t]);
            cpx<T> *dst = buf + ((j / Ns) * Ns * R + k) * BS + line;
#pragma unroll
            for (int r = 0; r < R; r++) dst[r * Ns * BS] = v[t][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int B, int BS, int NT, bool FWD, int S>
__device__ __forceinline__ void fft_stages_ip(cpx<T> *buf, const cpx<T> *tw) {
    constexpr RadixPlan P = make_radix_plan(L);
    if constexpr (S < P.n) {
        fft_stage_ip<T, L, P.r[S], P.ns[S], B, BS, NT, FWD>(buf, tw);
        fft_stages_ip<T, L, B, BS, NT, FWD, S + 1>(buf, tw);
    }
}

// B lines, point n of line c at buf[n*BS + c], transformed in place.
// Caller syncs before the call; the call ends with a barrier.
template <typename T, int L, int B, int BS, int NT, bool FWD>
__device__ __forceinline__ void fft_cols_ip(cpx<T> *buf, const cpx<T> *tw) {
    static_assert(fft_size_ok(L), "FFT length must be 2^a 3^b");
    fft_stages_ip<T, L, B, BS, NT, FWD, 0>(buf, tw);
}

// Two Hermitian half-spectrum rows A (even), B (odd) staged row-major in s
// -> the complex line Z = A + iB at point n (imaginary parts of kz = 0 and
// kz = L/2 dropped, as any c2r does).
template <typename T, int L>
__device__ __forceinline__ cpx<T> hermitian_pair(const cpx<T> *s, int p, int n) {
    constexpr int NZH = L / 2 + 1;
    const int     k   = (n <= L / 2) ? n : L - n;
    cpx<T>        A = s[(2 * p) * NZH + k], B = s[(2 * p + 1) * NZH + k];
    if (k == 0 || k == L / 2) {
        A.y = 0;
        B.y = 0;
    }
    if (n > L / 2) {
        A = cconj(A);
        B = cconj(B);
    }
    return cmk(A.x - B.y, A.y + B.x);
}

// Threads per workgroup of the FFT pass kernels: 8 wavefronts for FP32
// tiles, 4 for FP64 (whose tiles have half as many lines).
template <typename T>
struct PassThreads {
    static constexpr int value = sizeof(T) == 4 ? 512 : 256;
};

}  // namespace kh
