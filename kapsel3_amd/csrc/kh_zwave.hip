// kh_zwave.hip -- z passes on the wavefront-scope FFT engine (power-of-two
// NZ from 8 to 512).
//
// Rows along z are contiguous in both layouts (real R[x][y][z], spectral
// S[x][y][kz]), so every lane reads its points straight from HBM in the
// engine's input layout and no staging is needed on the way in. Two real
// rows share one complex line (row 2p in the real part, row 2p+1 in the
// imaginary part); the forward passes separate them again through the
// line's LDS strip, where Z[k] and Z[L-k] meet:
//   A[k] = (Z[k] + conj Z[L-k]) / 2,  B[k] = (Z[k] - conj Z[L-k]) / 2i.
// Only modes kz < kzv are read (the rest are zero) or written.
#include <type_traits>

#include "kh_fused.h"
#include "kh_wave_fft.h"
#include "kh_zwave.h"

namespace kh {

// Lane / line bookkeeping shared by the kernels below.
template <typename T, int L, int E>
struct WaveLane {
    using Geo = WaveGeom<T, L, E>;
    int       g;     // lane within the line
    long long line;  // complex line = real rows 2*line, 2*line+1
    bool      okA, okB;
    int       pitch;  // spectral row pitch (complex elements)
    cpx<T>   *buf;    // this line's LDS strip
    __device__ __forceinline__ WaveLane(cpx<T> *strips, long long nrows, int pitch_) : pitch(pitch_) {
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, w = lane / Geo::G;
        g              = lane - w * Geo::G;
        line           = ((long long)blockIdx.x * (Geo::NT / 64) + wv) * Geo::W + w;
        okA            = 2 * line < nrows;
        okB            = 2 * line + 1 < nrows;
        buf            = strips + (wv * Geo::W + w) * Geo::LP;
    }
};

// Two half-spectrum rows -> the complex line Z = A + iB in the input layout
// of the inverse engine F (imaginary parts of kz = 0 and L/2 dropped).
template <typename T, int L, typename F>
__device__ __forceinline__ void load_pair(const cpx<T> *__restrict__ src, const WaveLane<T, L, F::E> &wl, int kzv,
                                          cpx<T> *v) {
    const cpx<T> *ra = src + (2 * wl.line) * wl.pitch;
    const cpx<T> *rb = ra + wl.pitch;
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int n = F::in_point(wl.g, i);
        const int k = n <= L / 2 ? n : L - n;
        cpx<T>    a = cmk((T)0, (T)0), b = a;
        if (k < kzv) {
            if (wl.okA) a = ra[k];
            if (wl.okB) b = rb[k];
        }
        if (k == 0 || k == L / 2) {
            a.y = 0;
            b.y = 0;
        }
        if (n > L / 2) {
            a = cconj(a);
            b = cconj(b);
        }
        v[i] = cmk(a.x - b.y, a.y + b.x);
    }
}

// Forward engine output (layout of F) -> half spectra of the two rows,
// modes k < kzv written.
template <typename T, int L, typename F>
__device__ __forceinline__ void store_pair(const cpx<T> *v, const WaveLane<T, L, F::E> &wl, int kzv,
                                           cpx<T> *__restrict__ dst) {
    constexpr int NZH = L / 2 + 1, G = F::G;
#pragma unroll
    for (int i = 0; i < F::E; i++) wl.buf[WavePad<T>::at(F::out_point(wl.g, i))] = v[i];
    wave_sync();
    cpx<T> *ra = dst + (2 * wl.line) * wl.pitch;
    cpx<T> *rb = ra + wl.pitch;
#pragma unroll
    for (int t = 0; t < (NZH + G - 1) / G; t++) {
        const int k = wl.g + G * t;
        if (k < kzv) {
            const cpx<T> Zk = wl.buf[WavePad<T>::at(k)];
            const cpx<T> Zm = cconj(wl.buf[WavePad<T>::at((L - k) & (L - 1))]);
            const cpx<T> d  = csub(Zk, Zm);
            if (wl.okA) ra[k] = cscale(cadd(Zk, Zm), (T)0.5);
            if (wl.okB) rb[k] = cmk(d.y * (T)0.5, -d.x * (T)0.5);
        }
    }
    wave_sync();
}

// ------------------------------------------------------------- kernels
template <typename T, int L, int E>
__global__ __launch_bounds__(256) void k_zw_r2c(const T *__restrict__ in, cpx<T> *__restrict__ out, long long nrows,
                                                int pitch, int kzv) {
    using Geo = WaveGeom<T, L, E>;
    using F   = WaveFft<T, L, E, true, true>;
    __shared__ cpx<T> tw[F::TW], strips[(Geo::NT / 64) * Geo::W * Geo::LP];
    F::template fill_tw<Geo::NT>(tw);
    __syncthreads();
    const WaveLane<T, L, E> wl(strips, nrows, pitch);
    const T             *ra = in + (2 * wl.line) * L;
    const T             *rb = ra + L;
    cpx<T>               v[F::E];
#pragma unroll
    for (int i = 0; i < F::E; i++) {
        const int n = F::in_point(wl.g, i);
        v[i]        = cmk(wl.okA ? ra[n] : (T)0, wl.okB ? rb[n] : (T)0);
    }
    F::run(v, wl.buf, tw, wl.g);
    store_pair<T, L, F>(v, wl, kzv, out);
}

// Real rows of a line <-> its LDS strip as (row 2p, row 2p+1) pairs, four
// consecutive points per lane per access: each lane moves 16-byte vectors
// of both rows, so the HBM side of the z passes is fully coalesced.
template <typename T, int L>
struct RowStage {
    static constexpr int V  = 4;      // points per vector
    static constexpr int NV = L / V;  // vectors per row
    using vec = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    __device__ __forceinline__ static vec ld(const T *p) { return *reinterpret_cast<const vec *>(p); }
    __device__ __forceinline__ static void st(T *p, const vec &v) { *reinterpret_cast<vec *>(p) = v; }
};
template <typename V>
__device__ __forceinline__ bool nonzero(const V &v) {
    return !(v.x == 0 && v.y == 0 && v.z == 0 && v.w == 0);
}

template <typename T, int L, int E>
__global__ __launch_bounds__(256) void k_zw_c2r(ZIO<T> io, long long nrows, T scale, int kzv, int pitch) {
    using Geo = WaveGeom<T, L, E>;
    using F   = WaveFft<T, L, E, false, false>;
    using RS  = RowStage<T, L>;
    __shared__ cpx<T> tw[F::TW], strips[(Geo::NT / 64) * Geo::W * Geo::LP];
    F::template fill_tw<Geo::NT>(tw);
    __syncthreads();
    const WaveLane<T, L, E> wl(strips, nrows, pitch);
    cpx<T>                  v[F::E];
    load_pair<T, L, F>(io.in[blockIdx.y], wl, kzv, v);
    F::run(v, wl.buf, tw, wl.g);
#pragma unroll
    for (int i = 0; i < F::E; i++) wl.buf[WavePad<T>::at(F::out_point(wl.g, i))] = v[i];
    wave_sync();
    T *ra = io.out[blockIdx.y] + (2 * wl.line) * L;
    T *rb = ra + L;
    if constexpr (L >= 4 * Geo::G) {
#pragma unroll
        for (int t = 0; t < RS::NV / Geo::G; t++) {
            const int    z = 4 * (wl.g + Geo::G * t);
            const cpx<T> a = wl.buf[WavePad<T>::at(z)], b = wl.buf[WavePad<T>::at(z + 1)],
                         c = wl.buf[WavePad<T>::at(z + 2)], d = wl.buf[WavePad<T>::at(z + 3)];
            if (wl.okA) RS::st(ra + z, {a.x * scale, b.x * scale, c.x * scale, d.x * scale});
            if (wl.okB) RS::st(rb + z, {a.y * scale, b.y * scale, c.y * scale, d.y * scale});
        }
    } else {
#pragma unroll
        for (int i = 0; i < F::E; i++) {
            const int n = F::out_point(wl.g, i);
            if (wl.okA) ra[n] = v[i].x * scale;
            if (wl.okB) rb[n] = v[i].y * scale;
        }
    }
    wave_sync();
}

// The same inverse z pass for ncomp components in one workgroup, writing a
// real-space vector of 4 points only where the mask (phi_sum) has a nonzero
// point: u* is read only there afterwards (the F_H gather at points with
// phi > 0, the rigidity force where phi_sum != 0), so the rest of the field
// is left as it was. The mask is read once for all components; with an
// occupancy bitmask (occ.w) its bits replace the mask's values.
template <typename T, int L, int E>
__global__ __launch_bounds__(256) void k_zw_c2r_masked(ZIO<T> io, int ncomp, const T *__restrict__ mask,
                                                       long long nrows, T scale, int kzv, int pitch,
                                                       const int *__restrict__ skip, Grid bg, Occ occ) {
    using Geo = WaveGeom<T, L, E>;
    using F   = WaveFft<T, L, E, false, false>;
    using RS  = RowStage<T, L>;
    static_assert(L >= 4 * Geo::G, "four points per lane per row");
    constexpr int NT4 = RS::NV / Geo::G;  // vectors per lane per row
    if ((skip && *skip) || batch_skip(bg)) return;
    // io stays as passed (indexing a modified copy by the loop's c would
    // put it in scratch); each pointer is moved to this block's run at use
    mask                                  = brun(mask, bg);
    const unsigned long long *__restrict__ ow = occ_kept<T>() ? brun(occ.w, bg) : nullptr;
    __shared__ cpx<T> tw[F::TW], strips[(Geo::NT / 64) * Geo::W * Geo::LP];
    F::template fill_tw<Geo::NT>(tw);
    __syncthreads();
    const WaveLane<T, L, E> wl(strips, nrows, pitch);
    const long long         oa = (2 * wl.line) * L, ob = oa + L;
    bool                    wa[NT4], wb[NT4];
#pragma unroll
    for (int t = 0; t < NT4; t++) {
        const int z = 4 * (wl.g + Geo::G * t);
        if (ow) {
            wa[t] = wl.okA && occ_bit(ow, occ.words, 2 * wl.line, z >> 2);
            wb[t] = wl.okB && occ_bit(ow, occ.words, 2 * wl.line + 1, z >> 2);
        } else {
            wa[t] = wl.okA && nonzero(RS::ld(mask + oa + z));
            wb[t] = wl.okB && nonzero(RS::ld(mask + ob + z));
        }
    }
#pragma unroll 1
    for (int c = 0; c < ncomp; c++) {
        cpx<T> v[F::E];
        load_pair<T, L, F>(brun(io.in[c], bg), wl, kzv, v);
        F::run(v, wl.buf, tw, wl.g);
#pragma unroll
        for (int i = 0; i < F::E; i++) wl.buf[WavePad<T>::at(F::out_point(wl.g, i))] = v[i];
        wave_sync();
        T *const out = brun(io.out[c], bg);
        T       *ra = out + oa, *rb = out + ob;
#pragma unroll
        for (int t = 0; t < NT4; t++) {
            const int    z = 4 * (wl.g + Geo::G * t);
            const cpx<T> a = wl.buf[WavePad<T>::at(z)], b = wl.buf[WavePad<T>::at(z + 1)],
                         d = wl.buf[WavePad<T>::at(z + 2)], e = wl.buf[WavePad<T>::at(z + 3)];
            if (wa[t]) RS::st(ra + z, {a.x * scale, b.x * scale, d.x * scale, e.x * scale});
            if (wb[t]) RS::st(rb + z, {a.y * scale, b.y * scale, d.y * scale, e.y * scale});
        }
        wave_sync();
    }
}

// p, q, r = a x (x, y, z) per point of four-point vectors
template <typename V>
__device__ __forceinline__ void cross4(const V &a, const V &b, const V &c, const V &x, const V &y, const V &z, V &p,
                                       V &q, V &r) {
    p = {b.x * z.x - c.x * y.x, b.y * z.y - c.y * y.y, b.z * z.z - c.z * y.z, b.w * z.w - c.w * y.w};
    q = {c.x * x.x - a.x * z.x, c.y * x.y - a.y * z.y, c.z * x.z - a.z * z.z, c.w * x.w - a.w * z.w};
    r = {a.x * y.x - b.x * x.x, a.y * y.y - b.y * x.y, a.z * y.z - b.z * x.z, a.w * y.w - b.w * x.w};
}

// The inverse z pass of omega's three components fused with u x omega
// (contravariant u and omega, covariant product: operate_omega.cxx:113-137).
// A line's omega rows stay in registers, u is read once per point, and the
// product is written to io.out: omega itself never reaches HBM.
template <typename T, int L, int E>
__global__ __launch_bounds__(256) void k_zw_c2r_cross(ZIO<T> io, const T *__restrict__ u0, const T *__restrict__ u1,
                                                      const T *__restrict__ u2, long long nrows, T scale, int kzv,
                                                      int pitch) {
    using Geo = WaveGeom<T, L, E>;
    using F   = WaveFft<T, L, E, false, false>;
    using RS  = RowStage<T, L>;
    using V   = typename RS::vec;
    static_assert(L >= 4 * Geo::G, "four points per lane per row");
    constexpr int NT4 = RS::NV / Geo::G;  // vectors per lane per row
    __shared__ cpx<T> tw[F::TW], strips[(Geo::NT / 64) * Geo::W * Geo::LP];
    F::template fill_tw<Geo::NT>(tw);
    __syncthreads();
    const WaveLane<T, L, E> wl(strips, nrows, pitch);
    const long long         oa = (2 * wl.line) * L, ob = oa + L;
    V                       wa[3][NT4], wb[3][NT4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cpx<T> v[F::E];
        load_pair<T, L, F>(io.in[c], wl, kzv, v);
        F::run(v, wl.buf, tw, wl.g);
#pragma unroll
        for (int i = 0; i < F::E; i++) wl.buf[WavePad<T>::at(F::out_point(wl.g, i))] = v[i];
        wave_sync();
#pragma unroll
        for (int t = 0; t < NT4; t++) {
            const int    z = 4 * (wl.g + Geo::G * t);
            const cpx<T> a = wl.buf[WavePad<T>::at(z)], b = wl.buf[WavePad<T>::at(z + 1)],
                         d = wl.buf[WavePad<T>::at(z + 2)], e = wl.buf[WavePad<T>::at(z + 3)];
            wa[c][t] = {a.x * scale, b.x * scale, d.x * scale, e.x * scale};
            wb[c][t] = {a.y * scale, b.y * scale, d.y * scale, e.y * scale};
        }
        wave_sync();
    }
#pragma unroll
    for (int t = 0; t < NT4; t++) {
        const int z = 4 * (wl.g + Geo::G * t);
        V         p, q, r;
        if (wl.okA) {
            const long long o = oa + z;
            cross4(RS::ld(u0 + o), RS::ld(u1 + o), RS::ld(u2 + o), wa[0][t], wa[1][t], wa[2][t], p, q, r);
            RS::st(io.out[0] + o, p);
            RS::st(io.out[1] + o, q);
            RS::st(io.out[2] + o, r);
        }
        if (wl.okB) {
            const long long o = ob + z;
            cross4(RS::ld(u0 + o), RS::ld(u1 + o), RS::ld(u2 + o), wb[0][t], wb[1][t], wb[2][t], p, q, r);
            RS::st(io.out[0] + o, p);
            RS::st(io.out[1] + o, q);
            RS::st(io.out[2] + o, r);
        }
    }
}

// Advection products (U2advection_k, operate_omega.cxx:31-47): inverse z
// transforms of u1 (s0), u2 (s1), u3 (s2), the five products in registers,
// forward z transforms into p0..p4. With LIVE3 all three components stay in
// registers and each is read once; otherwise two are live at a time, so u2
// is read and transformed twice; the order is u1 u2 -> p2, then u3 -> p1, p4,
// then u2 again -> p0, p3. Either way a line's inputs are read before its
// products are written, so a product buffer may be an input buffer:
// p2 = s0, p1 = s2, p0 = s1 are allowed.
template <typename T, int L, int E, bool LIVE3>
__global__ __launch_bounds__(256) void k_zw_adv(Grid g, const cpx<T> *__restrict__ s0, const cpx<T> *__restrict__ s1,
                                                const cpx<T> *__restrict__ s2, cpx<T> *__restrict__ p0,
                                                cpx<T> *__restrict__ p1, cpx<T> *__restrict__ p2,
                                                cpx<T> *__restrict__ p3, cpx<T> *__restrict__ p4, T scale) {
    using Geo = WaveGeom<T, L, E>;
    using Fi  = WaveFft<T, L, E, false, false>;
    using Ff  = WaveFft<T, L, E, true, true>;
    static_assert(Fi::RN == Ff::R0, "inverse output layout must be the forward input layout");
    if (batch_skip(g)) return;
    s0 = brun(s0, g), s1 = brun(s1, g), s2 = brun(s2, g);
    p0 = brun(p0, g), p1 = brun(p1, g), p2 = brun(p2, g), p3 = brun(p3, g), p4 = brun(p4, g);
    __shared__ cpx<T> twi[Fi::TW], twf[Ff::TW], strips[(Geo::NT / 64) * Geo::W * Geo::LP];
    Fi::template fill_tw<Geo::NT>(twi);
    Ff::template fill_tw<Geo::NT>(twf);
    __syncthreads();
    const WaveLane<T, L, E> wl(strips, (long long)g.nx * g.ny, g.nzp);
    const int            kzv = g.trn_z;
    if constexpr (LIVE3) {
        cpx<T> a[E], b[E], c[E], v[E];
        load_pair<T, L, Fi>(s0, wl, kzv, a);
        load_pair<T, L, Fi>(s1, wl, kzv, b);
        load_pair<T, L, Fi>(s2, wl, kzv, c);
        Fi::run(a, wl.buf, twi, wl.g);
        Fi::run(b, wl.buf, twi, wl.g);
        Fi::run(c, wl.buf, twi, wl.g);
#pragma unroll
        for (int i = 0; i < E; i++) {
            a[i] = cscale(a[i], scale);
            b[i] = cscale(b[i], scale);
            c[i] = cscale(c[i], scale);
            v[i] = cmk(a[i].x * b[i].x, a[i].y * b[i].y);  // u1 u2
        }
        Ff::run(v, wl.buf, twf, wl.g);
        store_pair<T, L, Ff>(v, wl, kzv, p2);
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = cmk(c[i].x * a[i].x, c[i].y * a[i].y);  // u3 u1
        Ff::run(v, wl.buf, twf, wl.g);
        store_pair<T, L, Ff>(v, wl, kzv, p1);
#pragma unroll
        for (int i = 0; i < E; i++)  // u3^2 - u1^2
            v[i] = cmk((c[i].x - a[i].x) * (c[i].x + a[i].x), (c[i].y - a[i].y) * (c[i].y + a[i].y));
        Ff::run(v, wl.buf, twf, wl.g);
        store_pair<T, L, Ff>(v, wl, kzv, p4);
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = cmk(b[i].x * c[i].x, b[i].y * c[i].y);  // u2 u3
        Ff::run(v, wl.buf, twf, wl.g);
        store_pair<T, L, Ff>(v, wl, kzv, p0);
#pragma unroll
        for (int i = 0; i < E; i++)  // u2^2 - u3^2
            v[i] = cmk((b[i].x - c[i].x) * (b[i].x + c[i].x), (b[i].y - c[i].y) * (b[i].y + c[i].y));
        Ff::run(v, wl.buf, twf, wl.g);
        store_pair<T, L, Ff>(v, wl, kzv, p3);
        return;
    }
    cpx<T> ua[E], ub[E], v[E];
    load_pair<T, L, Fi>(s0, wl, kzv, ua);
    Fi::run(ua, wl.buf, twi, wl.g);
    load_pair<T, L, Fi>(s1, wl, kzv, ub);
    Fi::run(ub, wl.buf, twi, wl.g);
#pragma unroll
    for (int i = 0; i < E; i++) {
        ua[i] = cscale(ua[i], scale);
        ub[i] = cscale(ub[i], scale);
        v[i]  = cmk(ua[i].x * ub[i].x, ua[i].y * ub[i].y);  // u1 u2
    }
    Ff::run(v, wl.buf, twf, wl.g);
    store_pair<T, L, Ff>(v, wl, kzv, p2);
    load_pair<T, L, Fi>(s2, wl, kzv, ub);  // ub := u3
    Fi::run(ub, wl.buf, twi, wl.g);
#pragma unroll
    for (int i = 0; i < E; i++) {
        ub[i] = cscale(ub[i], scale);
        v[i]  = cmk(ub[i].x * ua[i].x, ub[i].y * ua[i].y);  // u3 u1
    }
    Ff::run(v, wl.buf, twf, wl.g);
    store_pair<T, L, Ff>(v, wl, kzv, p1);
#pragma unroll
    for (int i = 0; i < E; i++)  // u3^2 - u1^2
        v[i] = cmk((ub[i].x - ua[i].x) * (ub[i].x + ua[i].x), (ub[i].y - ua[i].y) * (ub[i].y + ua[i].y));
    Ff::run(v, wl.buf, twf, wl.g);
    store_pair<T, L, Ff>(v, wl, kzv, p4);
    load_pair<T, L, Fi>(s1, wl, kzv, ua);  // ua := u2 again
    Fi::run(ua, wl.buf, twi, wl.g);
#pragma unroll
    for (int i = 0; i < E; i++) {
        ua[i] = cscale(ua[i], scale);
        v[i]  = cmk(ua[i].x * ub[i].x, ua[i].y * ub[i].y);  // u2 u3
    }
    Ff::run(v, wl.buf, twf, wl.g);
    store_pair<T, L, Ff>(v, wl, kzv, p0);
#pragma unroll
    for (int i = 0; i < E; i++)  // u2^2 - u3^2
        v[i] = cmk((ua[i].x - ub[i].x) * (ua[i].x + ub[i].x), (ua[i].y - ub[i].y) * (ua[i].y + ub[i].y));
    Ff::run(v, wl.buf, twf, wl.g);
    store_pair<T, L, Ff>(v, wl, kzv, p3);
}

// Rigidity force f = up - min(phi_sum, 1) u* (Make_f_particle_dt_nonsole,
// f_particle.h:35-54; make_phi.cxx:248-259), forward z per component. With
// us2, u* = us + us2 (the squirmer step's u after the slip iteration, formed
// here as the reference forms it, u + up, sp_3d_ns.cxx:146-150); with no us
// and no phis, f = up read at every point (the slip force field of the slip
// iteration, nonzero on the whole shell |r - a| < xi/2, also where phi has
// underflowed to 0). clear: bit 0 leaves up zero, bit 1 phi_sum (with
// phis). A nonzero *skip makes the launch do nothing.
template <typename T, int L, int E>
__global__ __launch_bounds__(256) void k_zw_force(Grid g, ZForceIO<T> io, int clear, int *__restrict__ tick,
                                                  const int *__restrict__ skip) {
    using Geo = WaveGeom<T, L, E>;
    using Ff  = WaveFft<T, L, E, true, true>;
    using RS  = RowStage<T, L>;
    static_assert(L >= 4 * Geo::G, "four points per lane per row");
    constexpr int NT4 = RS::NV / Geo::G;  // vectors per lane per row
    if ((skip && *skip) || batch_skip(g)) return;
    // io stays as passed (a modified copy indexed by the loop's d would sit
    // in scratch); each pointer is moved to this block's run at use
    tick = brun(tick, g);
    __shared__ cpx<T> tw[Ff::TW], strips[(Geo::NT / 64) * Geo::W * Geo::LP];
    Ff::template fill_tw<Geo::NT>(tw);
    __syncthreads();
    if (tick && blockIdx.x == 0 && threadIdx.x == 0) *tick += 1;  // the step counter, after every reader
    const long long         nrows = (long long)g.nx * g.ny;
    const WaveLane<T, L, E> wl(strips, nrows, g.nzp);
    const long long         oa = (2 * wl.line) * L, ob = oa + L;
    using vec                  = typename RS::vec;
    const vec zero             = {(T)0, (T)0, (T)0, (T)0};
    T        *phis             = brun(io.phis, g);
    unsigned long long *const ow = occ_kept<T>() ? brun(io.occ.w, g) : nullptr;
    // Outside every particle phi_sum = 0, and so is u_p (it is scattered on
    // the same points with the same phi): f = 0 there whatever u* is, so
    // u_p and u* are read only where some phi_sum of the vector is nonzero.
    vec  pa[NT4], pb[NT4];  // min(phi_sum, 1) of both rows
    bool na[NT4], nb[NT4];
#pragma unroll
    for (int t = 0; t < NT4; t++) {
        if (ow) {  // the occupancy bits: f = up (u* absent), phi_sum only cleared
            const int z = 4 * (wl.g + Geo::G * t);
            na[t]       = wl.okA && occ_bit(ow, io.occ.words, 2 * wl.line, z >> 2);
            nb[t]       = wl.okB && occ_bit(ow, io.occ.words, 2 * wl.line + 1, z >> 2);
            pa[t] = pb[t] = zero;
            if ((clear & 2) && phis) {
                if (na[t]) RS::st(phis + oa + z, zero);
                if (nb[t]) RS::st(phis + ob + z, zero);
            }
            continue;
        }
        if (!phis) {  // no mask: up is read everywhere (f = up)
            na[t] = wl.okA;
            nb[t] = wl.okB;
            pa[t] = pb[t] = zero;
            continue;
        }
        const int z = 4 * (wl.g + Geo::G * t);
        vec       a = wl.okA ? RS::ld(phis + oa + z) : zero, b = wl.okB ? RS::ld(phis + ob + z) : zero;
        na[t]       = nonzero(a);
        nb[t]       = nonzero(b);
        if (clear & 2) {  // leave phi_sum zero for the next step's scatter
            if (na[t]) RS::st(phis + oa + z, zero);
            if (nb[t]) RS::st(phis + ob + z, zero);
        }
        pa[t]       = {fmin(a.x, (T)1), fmin(a.y, (T)1), fmin(a.z, (T)1), fmin(a.w, (T)1)};
        pb[t]       = {fmin(b.x, (T)1), fmin(b.y, (T)1), fmin(b.z, (T)1), fmin(b.w, (T)1)};
    }
#pragma unroll 1
    for (int d = 0; d < 3; d++) {
        T *const       UP  = brun(io.up[d], g);
        const T *const US  = brun(io.us[d], g), *const US2 = brun(io.us2[d], g);
        auto     ld_u = [&](long long o) -> vec {
            if (!US) return zero;
            const vec a = RS::ld(US + o);
            if (!US2) return a;
            const vec b = RS::ld(US2 + o);
            return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
        };
#pragma unroll
        for (int t = 0; t < NT4; t++) {
            const int z  = 4 * (wl.g + Geo::G * t);
            const vec ua = na[t] ? RS::ld(UP + oa + z) : zero, ub = nb[t] ? RS::ld(UP + ob + z) : zero;
            const vec sa = na[t] ? ld_u(oa + z) : zero, sb = nb[t] ? ld_u(ob + z) : zero;
            if (clear & 1) {  // and u_p zero
                if (na[t] && nonzero(ua)) RS::st(UP + oa + z, zero);
                if (nb[t] && nonzero(ub)) RS::st(UP + ob + z, zero);
            }
            wl.buf[WavePad<T>::at(z)]     = cmk(ua.x - pa[t].x * sa.x, ub.x - pb[t].x * sb.x);
            wl.buf[WavePad<T>::at(z + 1)] = cmk(ua.y - pa[t].y * sa.y, ub.y - pb[t].y * sb.y);
            wl.buf[WavePad<T>::at(z + 2)] = cmk(ua.z - pa[t].z * sa.z, ub.z - pb[t].z * sb.z);
            wl.buf[WavePad<T>::at(z + 3)] = cmk(ua.w - pa[t].w * sa.w, ub.w - pb[t].w * sb.w);
        }
        wave_sync();
        cpx<T> v[E];
#pragma unroll
        for (int i = 0; i < E; i++) v[i] = wl.buf[WavePad<T>::at(Ff::in_point(wl.g, i))];
        wave_sync();
        Ff::run(v, wl.buf, tw, wl.g);
        store_pair<T, L, Ff>(v, wl, g.trn_z, brun(io.f[d], g));
    }
    if (ow && (clear & 2)) {  // leave the bits zero for the next step's scatter
#pragma unroll
        for (int t = 0; t < NT4; t++) {
            const int v = wl.g + Geo::G * t;
            if ((v & 63) == 0) {
                if (wl.okA) ow[(2 * wl.line) * io.occ.words + (v >> 6)] = 0ull;
                if (wl.okB) ow[(2 * wl.line + 1) * io.occ.words + (v >> 6)] = 0ull;
            }
        }
    }
}

// ------------------------------------------------------------ launchers
#define KH_FOR_EACH_WL(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512)

bool zw_supported(int L) { return wave_fft_ok(L); }

// points per lane of each pass (the advection pass holds three components
// in registers, so it may take fewer points per lane; the force pass, which
// streams several fields per point, runs faster at four in both precisions)
#ifndef KH_ADV_E32
#define KH_ADV_E32 8
#endif
#ifndef KH_PASS_E32
#define KH_PASS_E32 8
#endif
#ifndef KH_ADV_E64
#define KH_ADV_E64 8
#endif
#ifndef KH_PASS_E64
#define KH_PASS_E64 8
#endif
#ifndef KH_FORCE_E64
#define KH_FORCE_E64 4
#endif
#ifndef KH_FORCE_E32
#define KH_FORCE_E32 4
#endif
// the advection pass keeps all three components live (each read once) at
// this many points per lane or fewer: at 8 (FP32 and FP64) the extra
// registers cost less than the second read and transform of u2
#ifndef KH_ADV_LIVE3_E
#define KH_ADV_LIVE3_E 8
#endif
template <typename T, int L>
struct ZwE {
    static constexpr int  pass  = WaveE<T, L, sizeof(T) == 4 ? KH_PASS_E32 : KH_PASS_E64>::value;
    static constexpr int  adv   = WaveE<T, L, sizeof(T) == 4 ? KH_ADV_E32 : KH_ADV_E64>::value;
    static constexpr bool live3 = adv <= KH_ADV_LIVE3_E;
    static constexpr int  force = WaveE<T, L, sizeof(T) == 4 ? KH_FORCE_E32 : KH_FORCE_E64>::value;
};

template <typename T, int L, int E>
static inline unsigned zw_blocks(long long nrows) {
    const long long lines = (nrows + 1) / 2;
    return (unsigned)((lines + WaveGeom<T, L, E>::LINES - 1) / WaveGeom<T, L, E>::LINES);
}

template <typename T>
hipError_t zw_r2c(int L, const T *in, cpx<T> *out, long long nrows, int pitch, hipStream_t s, int kzv) {
    if (kzv < 0 || kzv > L / 2 + 1) kzv = L / 2 + 1;
    switch (L) {
#define KH_CASE(n)                                                                                   \
    case n:                                                                                          \
        hipLaunchKernelGGL((k_zw_r2c<T, n, ZwE<T, n>::pass>), dim3(zw_blocks<T, n, ZwE<T, n>::pass>(nrows)), dim3(256), 0, s, in, out, nrows, pitch, kzv); \
        return hipGetLastError();
        KH_FOR_EACH_WL(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t zw_c2r(int L, const ZIO<T> &io, int ncomp, long long nrows, T scale, int kzv, int pitch,
                  hipStream_t s, const T *mask, const int *skip, const Grid *batch, Occ occ) {
    const Grid bg     = batch ? *batch : Grid{};
    const bool masked = (mask || occ.w) && L >= 32;
    if (occ.w && !masked) return hipErrorInvalidValue;
    if (batch_z(bg) > 1 && !masked) return hipErrorInvalidValue;  // only the masked pass is batched
    if (skip && !masked) return hipErrorInvalidValue;  // only the masked pass takes a skip flag
    if (masked) {  // every length with 4 points per lane per row
        switch (L) {
#define KH_CASE(n)                                                                                              \
    case n:                                                                                                     \
        if constexpr (n >= 32) {                                                                                \
            hipLaunchKernelGGL((k_zw_c2r_masked<T, n, ZwE<T, n>::pass>), dim3(zw_blocks<T, n, ZwE<T, n>::pass>(nrows), 1, batch_z(bg)), dim3(256), 0, s, io, ncomp, mask, nrows, scale, kzv, pitch, skip, bg, occ); \
            return hipGetLastError();                                                                           \
        }                                                                                                       \
        break;
            KH_FOR_EACH_WL(KH_CASE)
#undef KH_CASE
            default: return hipErrorInvalidValue;
        }
    }
    switch (L) {
#define KH_CASE(n)                                                                                              \
    case n:                                                                                                     \
        hipLaunchKernelGGL((k_zw_c2r<T, n, ZwE<T, n>::pass>), dim3(zw_blocks<T, n, ZwE<T, n>::pass>(nrows), ncomp), dim3(256), 0, s, io, nrows, scale, kzv, pitch); \
        return hipGetLastError();
        KH_FOR_EACH_WL(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

bool zw_cross_supported(int L) { return zw_supported(L) && L >= 32; }

template <typename T>
hipError_t zw_c2r_cross(int L, const ZIO<T> &io, const T *const u[3], long long nrows, T scale, int kzv, int pitch,
                        hipStream_t s) {
    switch (L) {
#define KH_CASE(n)                                                                                              \
    case n:                                                                                                     \
        if constexpr (n >= 32) {                                                                                \
            hipLaunchKernelGGL((k_zw_c2r_cross<T, n, ZwE<T, n>::pass>), dim3(zw_blocks<T, n, ZwE<T, n>::pass>(nrows)), dim3(256), 0, s, io, u[0], u[1], u[2], nrows, scale, kzv, pitch); \
            return hipGetLastError();                                                                           \
        }                                                                                                       \
        break;
        KH_FOR_EACH_WL(KH_CASE)
#undef KH_CASE
        default: break;
    }
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t zw_adv(const Grid &g, cpx<T> *const s3[3], cpx<T> *const p5[5], T scale, hipStream_t s) {
    const long long nrows = (long long)g.nx * g.ny;
    switch (g.nz) {
#define KH_CASE(n)                                                                                                \
    case n:                                                                                                       \
        hipLaunchKernelGGL((k_zw_adv<T, n, ZwE<T, n>::adv, ZwE<T, n>::live3>), dim3(zw_blocks<T, n, ZwE<T, n>::adv>(nrows), 1, batch_z(g)), dim3(256), 0, s, g, s3[0], s3[1], s3[2], \
                           p5[0], p5[1], p5[2], p5[3], p5[4], scale);                                             \
        return hipGetLastError();
        KH_FOR_EACH_WL(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t zw_force(const Grid &g, T *const up[3], T *phis, T *const us[3], cpx<T> *const f[3], int clear,
                    int *tick, hipStream_t s, T *const us2[3], const int *skip, Occ occ) {
    const long long nrows = (long long)g.nx * g.ny;
    if (occ.w && (us || us2)) return hipErrorInvalidValue;  // the bits replace phi_sum only where f = up
    ZForceIO<T>     io{};
    io.occ = occ;
    for (int d = 0; d < 3; d++) {
        io.up[d]  = up[d];
        io.us[d]  = us ? us[d] : nullptr;
        io.us2[d] = us2 ? us2[d] : nullptr;
        io.f[d]   = f[d];
    }
    io.phis = phis;
    switch (g.nz) {
#define KH_CASE(n)                                                                                                  \
    case n:                                                                                                         \
        hipLaunchKernelGGL((k_zw_force<T, n, ZwE<T, n>::force>), dim3(zw_blocks<T, n, ZwE<T, n>::force>(nrows), 1, batch_z(g)), dim3(256), 0, s, g, io, \
                           clear, tick, skip);                                                                      \
        return hipGetLastError();
        KH_FOR_EACH_WL(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

#define KH_INST(T)                                                                                           \
    template hipError_t zw_r2c<T>(int, const T *, cpx<T> *, long long, int, hipStream_t, int);               \
    template hipError_t zw_c2r<T>(int, const ZIO<T> &, int, long long, T, int, int, hipStream_t, const T *, \
                                  const int *, const Grid *, Occ);                                         \
    template hipError_t zw_c2r_cross<T>(int, const ZIO<T> &, const T *const[3], long long, T, int, int, hipStream_t); \
    template hipError_t zw_adv<T>(const Grid &, cpx<T> *const[3], cpx<T> *const[5], T, hipStream_t);         \
    template hipError_t zw_force<T>(const Grid &, T *const[3], T *, T *const[3], cpx<T> *const[3], int, int *,        \
                                    hipStream_t, T *const[3], const int *, Occ);
KH_INST(double)
KH_INST(float)
#undef KH_INST

}  // namespace kh
