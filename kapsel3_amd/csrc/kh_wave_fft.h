// kh_wave_fft.h -- wavefront-scope FFT of whole lines held in registers.
//
// A line of L complex points (L a power of two, 8..512) is owned by G = L/E
// lanes of one wavefront, E points per lane; a 64-lane wavefront holds
// W = 64/G lines. The transform is a Stockham decomposition L = R_0 R_1 ...
// with every radix dividing E: each stage does E/R radix-R butterflies in
// registers, and the exchange between two stages goes through a per-line
// LDS strip that only the owning wavefront touches, so no workgroup barrier
// is needed anywhere (a wavefront executes its LDS operations in order;
// wave_sync() only stops the compiler from moving them across each other).
//
// Layouts (lane g of the line, register i):
//   input  : i = b*R_0 + r  <->  point (g + G b) + r L/R_0
//   output : i = b*R_n + r  <->  point (g + G b) + r L/R_n   (R_n last radix)
// An inverse plan followed by the reversed forward plan therefore chains in
// registers: the inverse's output layout is the forward's input layout.
#pragma once

#include "kh_fft_engine.h"

namespace kh {

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// cos(pi m / 8), sin(pi m / 8)
__host__ __device__ constexpr double c16(int m) {
    constexpr double c1 = 0.92387953251128675612818318939678829, s1 = 0.38268343236508977172845998403039887,
                     h  = 0.70710678118654752440084436210484904;
    constexpr double t[16] = {1, c1, h, s1, 0, -s1, -h, -c1, -1, -c1, -h, -s1, 0, s1, h, c1};
    return t[m & 15];
}
__host__ __device__ constexpr double s16(int m) { return c16(m + 12); }  // sin x = cos(x - pi/2)

// v * exp(-+ 2 pi i m / 16) for a compile-time m
template <typename T, int M, bool FWD>
__device__ __forceinline__ cpx<T> tw16(cpx<T> v) {
    constexpr int m = M & 15;
    if constexpr (m == 0) {
        return v;
    } else if constexpr (m == 4) {
        return cmul_mi<T, FWD>(v);
    } else if constexpr (m == 8) {
        return cmk(-v.x, -v.y);
    } else if constexpr (m == 12) {
        return cmul_mi<T, !FWD>(v);
    } else {
        constexpr T c = (T)c16(m), s = (T)(FWD ? -s16(m) : s16(m));
        return cmk(v.x * c - v.y * s, v.x * s + v.y * c);
    }
}

template <typename T, bool FWD>
struct Dft<T, 16, FWD> {
    // n = n1 + 4 n2, k = k1 + 4 k2: DFT4 over n2, twiddle W16^(n1 k1), DFT4 over n1
    template <int N1>
    __device__ __forceinline__ static void col(cpx<T> *v, cpx<T> (&y)[4][4]) {
        cpx<T> a[4] = {v[N1], v[N1 + 4], v[N1 + 8], v[N1 + 12]};
        Dft<T, 4, FWD>::run(a);
        y[N1][0] = a[0];
        y[N1][1] = tw16<T, N1 * 1, FWD>(a[1]);
        y[N1][2] = tw16<T, N1 * 2, FWD>(a[2]);
        y[N1][3] = tw16<T, N1 * 3, FWD>(a[3]);
    }
    __device__ __forceinline__ static void run(cpx<T> *v) {
        cpx<T> y[4][4];
        col<0>(v, y);
        col<1>(v, y);
        col<2>(v, y);
        col<3>(v, y);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            cpx<T> b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
            Dft<T, 4, FWD>::run(b);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = b[k2];
        }
    }
};

struct WavePlan {
    int n;
    int r[8];
    int ns[8];
};

// L = E^a * rem with rem | E; REV reverses the radix order.
__host__ __device__ constexpr WavePlan make_wave_plan(int L, int E, bool rev) {
    WavePlan p{};
    int      rs[8] = {};
    int      k = 0, m = L;
    while (m >= E && m % E == 0) {
        rs[k++] = E;
        m /= E;
    }
    if (m > 1) rs[k++] = m;
    p.n = k;
    for (int i = 0; i < k; i++) p.r[i] = rev ? rs[k - 1 - i] : rs[i];
    int acc = 1;
    for (int i = 0; i < k; i++) {
        p.ns[i] = acc;
        acc *= p.r[i];
    }
    return p;
}

// Twiddle table of a plan, one block per stage S >= 1: entry
// off(S) + (r-1) Ns + k holds W_L^(r k L/(Ns R)), r = 1..R-1, k = 0..Ns-1.
// The lanes of a stage read consecutive k, so the reads are free of LDS
// bank conflicts (a table indexed by the exponent r k L/(Ns R) is read at
// a stride of L/(Ns R) entries, up to 8-way conflicting).
__host__ __device__ constexpr int wave_tw_off(const WavePlan &p, int s) {
    int o = 0;
    for (int i = 1; i < s; i++) o += (p.r[i] - 1) * p.ns[i];
    return o;
}
__host__ __device__ constexpr int wave_tw_size(const WavePlan &p) { return wave_tw_off(p, p.n); }

// Points per lane: `pref` (16 for FP32, 8 for FP64 by default), at most L
// and at least L/64 (a line fits one wavefront).
template <typename T, int L, int PREF = (sizeof(T) == 4 ? 16 : 8)>
struct WaveE {
    static constexpr int value = L < PREF ? L : (L / PREF > 64 ? L / 64 : PREF);
};

__host__ __device__ constexpr bool wave_fft_ok(int L) { return L >= 8 && L <= 512 && (L & (L - 1)) == 0; }

// LDS strip of one line in 128-byte groups (8 FP64 / 16 FP32 complex).
// KH_WAVE_SWIZZLE 0: each group padded by one element; 1: no padding, the
// element's slot within its group XORed with the group number, so that
// both consecutive points and points a group apart fall in distinct banks.
#ifndef KH_WAVE_SWIZZLE
#define KH_WAVE_SWIZZLE 0
#endif
template <typename T>
struct WavePad {
    static constexpr int shift = sizeof(T) == 4 ? 4 : 3;  // 16 / 8 complex per 128 B
    __device__ __forceinline__ static int at(int p) {
        if constexpr (KH_WAVE_SWIZZLE) return p ^ ((p >> shift) & ((1 << shift) - 1));
        return p + (p >> shift);
    }
};
#ifndef KH_WAVE_STAGGER
#define KH_WAVE_STAGGER 1
#endif
template <typename T, int L>
struct WaveStrip {
    // padded length, rounded so that neighbouring strips start 64 bytes apart
    // modulo 256 bytes
    static constexpr int raw   = L + (KH_WAVE_SWIZZLE ? 0 : (L >> WavePad<T>::shift));
    static constexpr int unit  = 64 / (int)sizeof(cpx<T>);
    static constexpr int value = ((raw + 4 * unit - 1) / (4 * unit)) * 4 * unit + KH_WAVE_STAGGER * unit;
};

template <typename T, int L, int E_, bool FWD, bool REV>
struct WaveFft {
    static constexpr int      E = E_;
    static constexpr int      G = L / E;
    static constexpr WavePlan P = make_wave_plan(L, E, REV);
    static constexpr int      R0 = P.r[0], RN = P.r[P.n - 1];
    static constexpr int      TW = wave_tw_size(P) > 0 ? wave_tw_size(P) : 1;  // twiddle entries (LDS)
    static_assert(G >= 1 && G <= 64, "line must fit one wavefront");

    // the plan's twiddle table (forward W_L^m = exp(-2 pi i m / L); the
    // inverse multiplies by the conjugate), by the workgroup's NT threads
    template <int NT>
    __device__ __forceinline__ static void fill_tw(cpx<T> *tw) {
        for (int e = threadIdx.x; e < TW; e += NT) {
            int m = 0;
#pragma unroll
            for (int s = 1; s < P.n; s++) {
                const int o = wave_tw_off(P, s), n = (P.r[s] - 1) * P.ns[s];
                if (e >= o && e < o + n) {
                    const int r = 1 + (e - o) / P.ns[s], k = (e - o) % P.ns[s];
                    m           = r * k * (L / (P.ns[s] * P.r[s]));
                }
            }
            double sn, cs;
            sincospi(2.0 * (double)m / (double)L, &sn, &cs);
            tw[e] = cmk((T)cs, (T)(-sn));
        }
    }

    template <int S>
    __device__ __forceinline__ static void stages(cpx<T> *v, cpx<T> *buf, const cpx<T> *tw, int g) {
        constexpr int R = P.r[S], Ns = P.ns[S], NB = E / R, to = wave_tw_off(P, S);
        static_assert(E % R == 0, "radix must divide the points per lane");
#pragma unroll
        for (int b = 0; b < NB; b++) {
            const int j = g + G * b;
            const int k = j & (Ns - 1);
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const cpx<T> w = tw[to + (r - 1) * Ns + k];
                    v[b * R + r]   = FWD ? cmul(v[b * R + r], w) : cmulc(v[b * R + r], w);
                }
            }
            Dft<T, R, FWD>::run(v + b * R);
        }
        if constexpr (S + 1 < P.n) {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int j    = g + G * b;
                const int k    = j & (Ns - 1);
                const int base = (j / Ns) * Ns * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) buf[WavePad<T>::at(base + r * Ns)] = v[b * R + r];
            }
            wave_sync();
            constexpr int R2 = P.r[S + 1], M2 = L / R2, NB2 = E / R2;
#pragma unroll
            for (int b = 0; b < NB2; b++) {
                const int j = g + G * b;
#pragma unroll
                for (int r = 0; r < R2; r++) v[b * R2 + r] = buf[WavePad<T>::at(j + r * M2)];
            }
            wave_sync();
            stages<S + 1>(v, buf, tw, g);
        }
    }

    // v: E points of lane g in the input layout -> output layout.
    __device__ __forceinline__ static void run(cpx<T> *v, cpx<T> *buf, const cpx<T> *tw, int g) {
        stages<0>(v, buf, tw, g);
    }
    __device__ __forceinline__ static int in_point(int g, int i) {
        return (g + G * (i / R0)) + (i % R0) * (L / R0);
    }
    __device__ __forceinline__ static int out_point(int g, int i) {
        return (g + G * (i / RN)) + (i % RN) * (L / RN);
    }
};

// Lane geometry of a wave pass: 4 wavefronts per workgroup, W lines each.
template <typename T, int L, int E_>
struct WaveGeom {
    static constexpr int NT = 256;
    static constexpr int E  = E_;
    static constexpr int G  = L / E;
    static constexpr int W  = 64 / G;
    static constexpr int LP = WaveStrip<T, L>::value;
    static constexpr int LINES = (NT / 64) * W;  // lines per workgroup
};

}  // namespace kh
