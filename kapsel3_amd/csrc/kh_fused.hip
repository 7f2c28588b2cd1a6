// kh_fused.hip -- the dealiasing-pruned, fused FFT passes of the SPM step.
//
// Every spectral field the step transforms is 2/3-truncated on one side of
// the transform (Truncate_two_third_rule, fft_wrapper.h:754-796): inverse
// transforms start from truncated spectra, forward transforms feed consumers
// that only read kept modes. So only the kept part of each pass is touched:
//   y passes : columns kz < TRN_Z only; inverse passes read only kept ky
//              rows (the rest are zero), forward passes write only kept ky
//   x passes : columns (ky, kz) both kept only; forward passes write only
//              kept kx
//   z passes : read/write only kz < TRN_Z
// Pointwise work between two passes on the same axis is fused into one
// kernel: the real-space quadratic products sit between the inverse and the
// forward z pass (k_z_adv), the construction of u_hat from zeta sits in
// front of the inverse x pass (k_x_zeta2u_inv), and the projection, the
// addition of u* and the curl sit behind the forward x pass (k_x_f_zeta).
#include "kh_fft_engine.h"
#include "kh_fused.h"
#include "kh_modes.h"
#include "kh_zwave.h"

namespace kh {

template <typename T>
struct ColTile;
template <>
struct ColTile<double> {
    static constexpr int C = 8;  // 8 x 16 B = 128 B per line element
};
template <>
struct ColTile<float> {
    static constexpr int C = 16;  // 16 x 8 B = 128 B
};

// Column tile and workgroup size of the fused x passes (all components in
// one workgroup)
template <typename T>
struct XTile {
    static constexpr int C  = ColTile<T>::C;
    static constexpr int NT = PassThreads<T>::value;
};
// zeta -> u_hat -> inverse x: FP32 runs narrower tiles on 4-wavefront
// workgroups (more workgroups resident, a shorter last round)
template <typename T>
struct XZTile {
    static constexpr int C  = sizeof(T) == 4 ? 8 : XTile<T>::C;
    static constexpr int NT = sizeof(T) == 4 ? 256 : XTile<T>::NT;
};

template <typename T>
struct ZTile;
template <>
struct ZTile<double> {
    static constexpr int ELEM = 2048;  // complex points per z block
};
template <>
struct ZTile<float> {
    static constexpr int ELEM = 2048;
};


// ------------------------------------------------------------------------
// Column passes. Y mode: outer block = x, line along y, columns kz < ncols.
// X mode: outer block = kept ky row, line along x, columns kz < ncols.
// in_mask: rows whose wavenumber is truncated are taken as zero (not read).
// out_mask: such rows are not written.
template <typename T, int L, bool FWD, bool XMODE>
__global__ __launch_bounds__(PassThreads<T>::value) void k_cols(Comps<T> comps, Grid g, int ncols, int in_mask,
                                                                int out_mask, const int *__restrict__ skip) {
    if ((skip && *skip) || batch_skip(g)) return;
    cpx<T> *__restrict__ data = brun(comps.p[blockIdx.y], g);  // one component per grid row
    constexpr int NT  = PassThreads<T>::value;
    constexpr int C   = ColTile<T>::C;
    constexpr int BS  = ColStride<T, C>::value;
    constexpr int NPT = (C * L + NT - 1) / NT;
    __shared__ cpx<T> s[L * BS], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const int ntiles = (ncols + C - 1) / C;
    const int o = blockIdx.x / ntiles, kz0 = (blockIdx.x - o * ntiles) * C;
    long long base, lstride;
    int       trn, nline;
    if (XMODE) {
        const int j = kept_index(o, g.ny, g.trn_y);
        base        = (long long)j * g.nzp + kz0;
        lstride     = (long long)g.ny * g.nzp;
        trn         = g.trn_x;
        nline       = g.nx;
    } else {
        base    = (long long)o * g.ny * g.nzp + kz0;
        lstride = g.nzp;
        trn     = g.trn_y;
        nline   = g.ny;
    }
    cpx<T> v[NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int e = threadIdx.x + i * NT;
        const int n = e / C, c = e - n * C;
        const int k = (n > nline / 2) ? n - nline : n;
        v[i]        = cmk((T)0, (T)0);
        if (e < C * L && kz0 + c < ncols && (!in_mask || kept1(k, trn))) v[i] = data[base + n * lstride + c];
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
        const int k = (n > nline / 2) ? n - nline : n;
        if (e < C * L && kz0 + c < ncols && (!out_mask || kept1(k, trn))) data[base + n * lstride + c] = s[n * BS + c];
    }
}

// ------------------------------------------------------------------------
// u_hat from zeta (Zeta_k2u_k, fft_wrapper.cxx:591-688, on the truncated
// zeta) for one mode; (0,0,0) gives uk_dc.
template <typename T>
__device__ __forceinline__ void zeta_to_uhat(const Grid &g, int ikx, int iky, int kz, cpx<T> d1, cpx<T> d2,
                                             const double *uk, cpx<T> *u) {
    if (ikx == 0 && iky == 0 && kz == 0) {
        u[0] = cmk((T)uk[0], (T)0);
        u[1] = cmk((T)uk[1], (T)0);
        u[2] = cmk((T)uk[2], (T)0);
        return;
    }
    const double kx = g.wave_x * ikx, ky = g.wave_y * iky, kz_ = g.wave_z * kz;
    const double k2 = kx * kx + ky * ky + kz_ * kz_;
    const double ik2 = (k2 > 0.0) ? 1. / k2 : 0.0;
    double       o0r, o0i, o1r, o1i, o2r, o2i;
    if (ikx != 0) {
        o0r = -(1. / kx) * (ky * (double)d1.x + kz_ * (double)d2.x);
        o0i = -(1. / kx) * (ky * (double)d1.y + kz_ * (double)d2.y);
        o1r = d1.x;
        o1i = d1.y;
        o2r = d2.x;
        o2i = d2.y;
    } else if (iky != 0) {
        const double dmy = -(kz_ / ky);
        o0r = d2.x;
        o0i = d2.y;
        o1r = dmy * (double)d1.x;
        o1i = dmy * (double)d1.y;
        o2r = d1.x;
        o2i = d1.y;
    } else {
        o0r = d1.x;
        o0i = d1.y;
        o1r = d2.x;
        o1i = d2.y;
        o2r = 0.0;
        o2i = 0.0;
    }
    const double qx = kx * ik2, qy = ky * ik2, qz = kz_ * ik2;
    u[0] = cmk((T)(-(qy * o2i - qz * o1i)), (T)(qy * o2r - qz * o1r));
    u[1] = cmk((T)(-(qz * o0i - qx * o2i)), (T)(qz * o0r - qx * o2r));
    u[2] = cmk((T)(-(qx * o1i - qy * o0i)), (T)(qx * o1r - qy * o0r));
}

// zeta -> u_hat (3 components) on the kept (ky,kz) columns, then the
// inverse x pass of each component. fixmask: FIX_CELL zeroing of uk_dc
// (sp_3d_ns.cxx:46-52), done by the block holding (ky,kz) = (0,0).
template <typename T, int L>
__global__ __launch_bounds__(XZTile<T>::NT) void k_x_zeta2u_inv(
    Grid g, const cpx<T> *__restrict__ z0, const cpx<T> *__restrict__ z1, cpx<T> *__restrict__ u0,
    cpx<T> *__restrict__ u1, cpx<T> *__restrict__ u2, double *__restrict__ ukdc, int fixmask) {
    if (batch_skip(g)) return;
    z0 = brun(z0, g), z1 = brun(z1, g), u0 = brun(u0, g), u1 = brun(u1, g), u2 = brun(u2, g);
    ukdc = brun(ukdc, g);
    constexpr int NT  = XZTile<T>::NT;
    constexpr int C   = XZTile<T>::C;
    constexpr int BS  = ColStride<T, C>::value;
    constexpr int NPT = (C * L + NT - 1) / NT;
    __shared__ cpx<T> s[L * BS], tw[L];
    __shared__ double uk[3];
    fill_twiddles<T, L, NT>(tw);
    const int ntiles = (g.trn_z + C - 1) / C;
    const int o = blockIdx.x / ntiles, kz0 = (blockIdx.x - o * ntiles) * C;
    const int j = kept_index(o, g.ny, g.trn_y), iky = (j > g.hny) ? j - g.ny : j;
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) uk[d] = (fixmask & (1 << d)) ? 0.0 : ukdc[d];
    }
    __syncthreads();
    if (fixmask && j == 0 && kz0 == 0 && threadIdx.x == 0)
        for (int d = 0; d < 3; d++)
            if (fixmask & (1 << d)) ukdc[d] = 0.0;
    const long long lstride = (long long)g.ny * g.nzp;
    const long long base    = (long long)j * g.nzp + kz0;
    cpx<T>          D1[NPT], D2[NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int p = threadIdx.x + i * NT;
        const int n = p / C, c = p - n * C, kz = kz0 + c;
        const int ikx = (n > g.hnx) ? n - g.nx : n;
        const bool ok = p < C * L && kz < g.trn_z && kept1(ikx, g.trn_x);
        D1[i]         = ok ? z0[base + n * lstride + c] : cmk((T)0, (T)0);
        D2[i]         = ok ? z1[base + n * lstride + c] : cmk((T)0, (T)0);
    }
    cpx<T> U[3][NPT];
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int p = threadIdx.x + i * NT;
        const int n = p / C, c = p - n * C, kz = kz0 + c;
        const int ikx = (n > g.hnx) ? n - g.nx : n;
        U[0][i] = U[1][i] = U[2][i] = cmk((T)0, (T)0);
        if (p < C * L && kz < g.trn_z && kept1(ikx, g.trn_x)) {
            cpx<T> u[3];
            zeta_to_uhat<T>(g, ikx, iky, kz, D1[i], D2[i], uk, u);
            U[0][i] = u[0];
            U[1][i] = u[1];
            U[2][i] = u[2];
        }
    }
    cpx<T> *out[3] = {u0, u1, u2};
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int p = threadIdx.x + i * NT;
            if (p < C * L) s[(p / C) * BS + p % C] = U[d][i];
        }
        __syncthreads();
        fft_cols_ip<T, L, C, BS, NT, false>(s, tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            const int n = e / C, c = e - n * C;
            if (e < C * L && kz0 + c < g.trn_z) out[d][base + n * lstride + c] = s[n * BS + c];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// z pass of the advection term: three half-spectrum rows per (x,y) (kz <
// TRN_Z nonzero) -> inverse z -> u in registers -> the five products of
// U2advection_k (operate_omega.cxx:31-47) -> forward z -> five half spectra
// (kz < TRN_Z written). Two real rows share one complex line.
template <typename T, int L>
__global__ __launch_bounds__(PassThreads<T>::value) void k_z_adv(
    Grid g, const cpx<T> *__restrict__ s0g, const cpx<T> *__restrict__ s1g, const cpx<T> *__restrict__ s2g,
    cpx<T> *__restrict__ p0, cpx<T> *__restrict__ p1, cpx<T> *__restrict__ p2, cpx<T> *__restrict__ p3,
    cpx<T> *__restrict__ p4, T scale) {
    if (batch_skip(g)) return;
    s0g = brun(s0g, g), s1g = brun(s1g, g), s2g = brun(s2g, g);
    p0 = brun(p0, g), p1 = brun(p1, g), p2 = brun(p2, g), p3 = brun(p3, g), p4 = brun(p4, g);
    constexpr int NT  = PassThreads<T>::value;
    constexpr int P   = (ZTile<T>::ELEM / L) > 0 ? ZTile<T>::ELEM / L : 1;
    constexpr int NZH = L / 2 + 1;
    constexpr int NPT = (P * L + NT - 1) / NT;
    constexpr int BS  = ColStride<T, P>::value;
    constexpr int SZ  = (L * BS > P * (L + 2)) ? L * BS : P * (L + 2);
    __shared__ cpx<T> s[SZ], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const long long nrows  = (long long)g.nx * g.ny;
    const long long row0   = (long long)blockIdx.x * (2 * P);
    const int       kzv    = g.trn_z;
    const cpx<T>   *src[3] = {s0g, s1g, s2g};
    cpx<T>          U[3][NPT];  // (row 2p, row 2p+1) of component d at z = n, point e = p*L + n
    for (int d = 0; d < 3; d++) {
        for (int e = threadIdx.x; e < 2 * P * NZH; e += NT) {
            const int rl = e / NZH, k = e - rl * NZH;
            s[e]         = (row0 + rl < nrows && k < kzv) ? src[d][(row0 + rl) * g.nzp + k] : cmk((T)0, (T)0);
        }
        __syncthreads();
        cpx<T> zv[NPT];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < P * L) zv[i] = hermitian_pair<T, L>(s, e / L, e % L);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < P * L) s[(e % L) * BS + e / L] = zv[i];
        }
        __syncthreads();
        fft_cols_ip<T, L, P, BS, NT, false>(s, tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            U[d][i]     = (e < P * L) ? cscale(s[(e % L) * BS + e / L], scale) : cmk((T)0, (T)0);
        }
        __syncthreads();
    }
    cpx<T> *dst[5] = {p0, p1, p2, p3, p4};
    for (int q = 0; q < 5; q++) {
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            if (e >= P * L) continue;
            cpx<T> v;
            // .x = row 2p, .y = row 2p+1
            const cpx<T> u1 = U[0][i], u2 = U[1][i], u3 = U[2][i];
            if (q == 0)
                v = cmk(u2.x * u3.x, u2.y * u3.y);
            else if (q == 1)
                v = cmk(u3.x * u1.x, u3.y * u1.y);
            else if (q == 2)
                v = cmk(u1.x * u2.x, u1.y * u2.y);
            else if (q == 3)
                v = cmk((u2.x - u3.x) * (u2.x + u3.x), (u2.y - u3.y) * (u2.y + u3.y));
            else
                v = cmk((u3.x - u1.x) * (u3.x + u1.x), (u3.y - u1.y) * (u3.y + u1.y));
            s[(e % L) * BS + e / L] = v;
        }
        __syncthreads();
        fft_cols_ip<T, L, P, BS, NT, true>(s, tw);
        for (int e = threadIdx.x; e < 2 * P * NZH; e += NT) {
            const int rl = e / NZH, k = e - rl * NZH;
            if (row0 + rl >= nrows) break;
            if (k >= kzv) continue;
            const int    p  = rl >> 1;
            const cpx<T> Zk = s[k * BS + p];
            const cpx<T> Zm = cconj(s[((L - k) % L) * BS + p]);
            cpx<T>       v;
            if (rl & 1) {
                const cpx<T> d = csub(Zk, Zm);
                v              = cmk(d.y * (T)0.5, -d.x * (T)0.5);
            } else {
                v = cscale(cadd(Zk, Zm), (T)0.5);
            }
            dst[q][(row0 + rl) * g.nzp + k] = v;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// z pass of the rigidity force: f = up - min(phi_sum, 1) u*
// (Make_f_particle_dt_nonsole, f_particle.h:35-54; make_phi.cxx:248-259),
// forward z per component, kz < TRN_Z written.
template <typename T, int L>
__global__ __launch_bounds__(PassThreads<T>::value) void k_z_force(
    Grid g, const T *__restrict__ up0, const T *__restrict__ up1, const T *__restrict__ up2,
    const T *__restrict__ phis, const T *__restrict__ us0, const T *__restrict__ us1, const T *__restrict__ us2,
    cpx<T> *__restrict__ f0, cpx<T> *__restrict__ f1, cpx<T> *__restrict__ f2) {
    if (batch_skip(g)) return;
    up0 = brun(up0, g), up1 = brun(up1, g), up2 = brun(up2, g), phis = brun(phis, g);
    us0 = brun(us0, g), us1 = brun(us1, g), us2 = brun(us2, g);
    f0 = brun(f0, g), f1 = brun(f1, g), f2 = brun(f2, g);
    constexpr int NT  = PassThreads<T>::value;
    constexpr int P   = (ZTile<T>::ELEM / L) > 0 ? ZTile<T>::ELEM / L : 1;
    constexpr int NZH = L / 2 + 1;
    constexpr int NR  = (2 * P * L + NT - 1) / NT;  // real points per thread
    constexpr int BS  = ColStride<T, P>::value;
    __shared__ cpx<T> s[L * BS], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const long long nrows = (long long)g.nx * g.ny;
    const long long row0  = (long long)blockIdx.x * (2 * P);
    const T        *UP[3] = {up0, up1, up2}, *US[3] = {us0, us1, us2};
    cpx<T>         *dst[3] = {f0, f1, f2};
    // real point e = rl*L + n of the block (contiguous rows)
    T ph[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) {
        const int e = threadIdx.x + i * NT;
        ph[i]       = (e < 2 * P * L && row0 + e / L < nrows) ? fmin(phis[row0 * L + e], (T)1) : (T)0;
    }
    for (int d = 0; d < 3; d++) {
        T a[NR], b[NR];
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int  e  = threadIdx.x + i * NT;
            const bool ok = e < 2 * P * L && row0 + e / L < nrows;
            a[i]          = ok ? UP[d][row0 * L + e] : (T)0;
            b[i]          = ok && US[d] ? US[d][row0 * L + e] : (T)0;
        }
#pragma unroll
        for (int i = 0; i < NR; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < 2 * P * L) {
                const int rl = e / L, n = e - rl * L;
                reinterpret_cast<T *>(&s[n * BS + (rl >> 1)])[rl & 1] = a[i] - ph[i] * b[i];
            }
        }
        __syncthreads();
        fft_cols_ip<T, L, P, BS, NT, true>(s, tw);
        for (int e = threadIdx.x; e < 2 * P * NZH; e += NT) {
            const int rl = e / NZH, k = e - rl * NZH;
            if (row0 + rl >= nrows) break;
            if (k >= g.trn_z) continue;
            const int    p  = rl >> 1;
            const cpx<T> Zk = s[k * BS + p];
            const cpx<T> Zm = cconj(s[((L - k) % L) * BS + p]);
            cpx<T>       v;
            if (rl & 1) {
                const cpx<T> dd = csub(Zk, Zm);
                v               = cmk(dd.y * (T)0.5, -dd.x * (T)0.5);
            } else {
                v = cscale(cadd(Zk, Zm), (T)0.5);
            }
            dst[d][(row0 + rl) * g.nzp + k] = v;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Forward x pass of the five advection products on the kept (ky,kz)
// columns, then per kept mode Omega_k2zeta_k + the slaved-Euler update of
// zeta (NS_solver_slavedEuler, fluid_solver.cxx:348-374): the products are
// never written back in spectral form.
template <typename T, int L>
__global__ __launch_bounds__(XTile<T>::NT) void k_x_adv_ns(Grid g, Comps<T> prod, cpx<T> *__restrict__ z0,
                                                                    cpx<T> *__restrict__ z1, double dmy0,
                                                                    double dmy1) {
    if (batch_skip(g)) return;
    for (int q = 0; q < 5; q++) prod.p[q] = brun(prod.p[q], g);
    z0 = brun(z0, g), z1 = brun(z1, g);
    constexpr int NT  = XTile<T>::NT;
    constexpr int C   = XTile<T>::C;
    constexpr int BS  = ColStride<T, C>::value;
    constexpr int NPT = (C * L + NT - 1) / NT;
    __shared__ cpx<T> s[L * BS], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const int ntiles = (g.trn_z + C - 1) / C;
    const int o = blockIdx.x / ntiles, kz0 = (blockIdx.x - o * ntiles) * C;
    const int j = kept_index(o, g.ny, g.trn_y), iky = (j > g.hny) ? j - g.ny : j;
    const long long lstride = (long long)g.ny * g.nzp;
    const long long base    = (long long)j * g.nzp + kz0;
    cpx<T>          Ph[5][NPT];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const cpx<T> *src = prod.p[q];
        cpx<T>        v[NPT];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            const int n = e / C, c = e - n * C;
            v[i]        = (e < C * L && kz0 + c < g.trn_z) ? src[base + n * lstride + c] : cmk((T)0, (T)0);
        }
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < C * L) s[(e / C) * BS + e % C] = v[i];
        }
        __syncthreads();
        fft_cols_ip<T, L, C, BS, NT, true>(s, tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int p = threadIdx.x + i * NT;
            if (p < C * L) Ph[q][i] = s[(p / C) * BS + p % C];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int p = threadIdx.x + i * NT;
        if (p >= C * L) continue;
        const int n = p / C, c = p - n * C, kz = kz0 + c;
        const int ikx = (n > g.hnx) ? n - g.nx : n;
        if (kz >= g.trn_z || !kept1(ikx, g.trn_x)) continue;
        const long long idx  = base + n * lstride + c;
        const cpx<T>    a[5] = {Ph[0][i], Ph[1][i], Ph[2][i], Ph[3][i], Ph[4][i]};
        cpx<T>          x0 = z0[idx], x1 = z1[idx];
        slaved_euler_mode<T>(g, ikx, iky, kz, a, dmy0, dmy1, x0, x1);
        z0[idx] = x0;
        z1[idx] = x1;
    }
}

// ------------------------------------------------------------------------
// Forward x pass of the rigidity force on the kept (ky,kz) columns, then per
// kept mode: truncation + Solenoidal_uk (operate_omega.cxx:171-195), the sum
// with u* rebuilt from zeta (the reference's U2zeta_k of u* + f), and
// U_k2zeta_k (fft_wrapper.cxx:407-476). Writes zeta at the kept modes and
// uk_dc += f_hat(0).
template <typename T, int L>
__global__ __launch_bounds__(XTile<T>::NT) void k_x_f_zeta(Grid g, cpx<T> *__restrict__ f0,
                                                                    cpx<T> *__restrict__ f1, cpx<T> *__restrict__ f2,
                                                                    cpx<T> *__restrict__ z0, cpx<T> *__restrict__ z1,
                                                                    double *__restrict__ ukdc) {
    if (batch_skip(g)) return;
    f0 = brun(f0, g), f1 = brun(f1, g), f2 = brun(f2, g), z0 = brun(z0, g), z1 = brun(z1, g), ukdc = brun(ukdc, g);
    constexpr int NT  = XTile<T>::NT;
    constexpr int C   = XTile<T>::C;
    constexpr int BS  = ColStride<T, C>::value;
    constexpr int NPT = (C * L + NT - 1) / NT;
    __shared__ cpx<T> s[L * BS], tw[L];
    __shared__ double uk[3];
    fill_twiddles<T, L, NT>(tw);
    const int ntiles = (g.trn_z + C - 1) / C;
    const int o = blockIdx.x / ntiles, kz0 = (blockIdx.x - o * ntiles) * C;
    const int j = kept_index(o, g.ny, g.trn_y), iky = (j > g.hny) ? j - g.ny : j;
    if (threadIdx.x < 3) uk[threadIdx.x] = ukdc[threadIdx.x];
    const long long lstride = (long long)g.ny * g.nzp;
    const long long base    = (long long)j * g.nzp + kz0;
    cpx<T>         *F[3]    = {f0, f1, f2};
    cpx<T>          Fh[3][NPT];
    for (int d = 0; d < 3; d++) {
        cpx<T> v[NPT];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            const int n = e / C, c = e - n * C;
            v[i]        = (e < C * L && kz0 + c < g.trn_z) ? F[d][base + n * lstride + c] : cmk((T)0, (T)0);
        }
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < C * L) s[(e / C) * BS + e % C] = v[i];
        }
        __syncthreads();
        fft_cols_ip<T, L, C, BS, NT, true>(s, tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int p = threadIdx.x + i * NT;
            if (p < C * L) Fh[d][i] = s[(p / C) * BS + p % C];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int p = threadIdx.x + i * NT;
        if (p >= C * L) continue;
        const int n = p / C, c = p - n * C, kz = kz0 + c;
        const int ikx = (n > g.hnx) ? n - g.nx : n;
        if (kz >= g.trn_z || !kept1(ikx, g.trn_x)) continue;
        const long long idx = base + n * lstride + c;
        const double    kx = ikx * g.wave_x, ky = iky * g.wave_y, kzz = kz * g.wave_z;
        const double    K2  = (g.wave_x * ikx) * (g.wave_x * ikx) + (g.wave_y * iky) * (g.wave_y * iky) +
                          (g.wave_z * kz) * (g.wave_z * kz);
        const double IK2 = (K2 > 0.0) ? 1. / K2 : 0.0;
        // Solenoidal_uk on f_hat
        const cpx<T> a = Fh[0][i], b = Fh[1][i], cc = Fh[2][i];
        const double dr = IK2 * ((double)a.x * kx + (double)b.x * ky + (double)cc.x * kzz);
        const double di = IK2 * ((double)a.y * kx + (double)b.y * ky + (double)cc.y * kzz);
        double       fr[3] = {a.x - dr * kx, b.x - dr * ky, cc.x - dr * kzz};
        double       fi[3] = {a.y - di * kx, b.y - di * ky, cc.y - di * kzz};
        // u* rebuilt from the (post-NS, truncated) zeta
        cpx<T> us[3];
        zeta_to_uhat<T>(g, ikx, iky, kz, z0[idx], z1[idx], uk, us);
        double ur[3], ui[3];
        for (int d = 0; d < 3; d++) {
            ur[d] = (double)us[d].x + fr[d];
            ui[d] = (double)us[d].y + fi[d];
        }
        if (ikx == 0 && iky == 0 && kz == 0) {  // k = 0 stays FP64: uk_dc + sum(f)
            for (int d = 0; d < 3; d++) ukdc[d] = uk[d] + fr[d];
        }
        double ore[3], oim[3];
        ore[0] = -(ky * ui[2] - kzz * ui[1]);
        oim[0] = (ky * ur[2] - kzz * ur[1]);
        ore[1] = -(kzz * ui[0] - kx * ui[2]);
        oim[1] = (kzz * ur[0] - kx * ur[2]);
        ore[2] = -(kx * ui[1] - ky * ui[0]);
        oim[2] = (kx * ur[1] - ky * ur[0]);
        int c0, c1;
        if (ikx != 0) {
            c0 = 1;
            c1 = 2;
        } else if (iky != 0) {
            c0 = 2;
            c1 = 0;
        } else {
            c0 = 0;
            c1 = 1;
        }
        z0[idx] = cmk((T)ore[c0], (T)oim[c0]);
        z1[idx] = cmk((T)ore[c1], (T)oim[c1]);
    }
}

// The slip iteration's projection (Solenoidal_u, operate_omega.h:105-112,
// which truncates to the 2/3 modes) less its z and y passes: forward x of
// the slip force field's three components on the kept (ky,kz) columns; per
// kept mode truncation + Solenoidal_uk (operate_omega.cxx:171-195) and the
// curl of the projected field (U_k2zeta_k, fft_wrapper.cxx:407-476) into
// (a0, a1), its k = 0 mode into dc; modes beyond the 2/3 rule set to zero;
// then the inverse x pass of each projected component, in place.
template <typename T, int L>
__global__ __launch_bounds__(XTile<T>::NT) void k_x_slip(Grid g, cpx<T> *__restrict__ f0, cpx<T> *__restrict__ f1,
                                                        cpx<T> *__restrict__ f2, cpx<T> *__restrict__ a0,
                                                        cpx<T> *__restrict__ a1, double *__restrict__ dc,
                                                        const int *__restrict__ skip) {
    constexpr int NT  = XTile<T>::NT;
    constexpr int C   = XTile<T>::C;
    constexpr int BS  = ColStride<T, C>::value;
    constexpr int NPT = (C * L + NT - 1) / NT;
    if (skip && *skip) return;
    __shared__ cpx<T> s[L * BS], tw[L];
    fill_twiddles<T, L, NT>(tw);
    const int ntiles = (g.trn_z + C - 1) / C;
    const int o = blockIdx.x / ntiles, kz0 = (blockIdx.x - o * ntiles) * C;
    const int j = kept_index(o, g.ny, g.trn_y), iky = (j > g.hny) ? j - g.ny : j;
    const long long lstride = (long long)g.ny * g.nzp;
    const long long base    = (long long)j * g.nzp + kz0;
    cpx<T>         *F[3]    = {f0, f1, f2};
    cpx<T>          Fh[3][NPT];
    for (int d = 0; d < 3; d++) {
        cpx<T> v[NPT];
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            const int n = e / C, c = e - n * C;
            v[i]        = (e < C * L && kz0 + c < g.trn_z) ? F[d][base + n * lstride + c] : cmk((T)0, (T)0);
        }
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < C * L) s[(e / C) * BS + e % C] = v[i];
        }
        __syncthreads();
        fft_cols_ip<T, L, C, BS, NT, true>(s, tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int p = threadIdx.x + i * NT;
            if (p < C * L) Fh[d][i] = s[(p / C) * BS + p % C];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < NPT; i++) {
        const int p = threadIdx.x + i * NT;
        if (p >= C * L) continue;
        const int n = p / C, c = p - n * C, kz = kz0 + c;
        const int ikx = (n > g.hnx) ? n - g.nx : n;
        if (kz >= g.trn_z) continue;
        if (!kept1(ikx, g.trn_x)) {
            Fh[0][i] = Fh[1][i] = Fh[2][i] = cmk((T)0, (T)0);
            continue;
        }
        const long long idx = base + n * lstride + c;
        const double    kx = ikx * g.wave_x, ky = iky * g.wave_y, kzz = kz * g.wave_z;
        const double    K2  = (g.wave_x * ikx) * (g.wave_x * ikx) + (g.wave_y * iky) * (g.wave_y * iky) +
                          (g.wave_z * kz) * (g.wave_z * kz);
        const double IK2 = (K2 > 0.0) ? 1. / K2 : 0.0;
        // Solenoidal_uk, stored in T as the reference's field is
        const cpx<T> a = Fh[0][i], b = Fh[1][i], cc = Fh[2][i];
        const double dr = IK2 * ((double)a.x * kx + (double)b.x * ky + (double)cc.x * kzz);
        const double di = IK2 * ((double)a.y * kx + (double)b.y * ky + (double)cc.y * kzz);
        const cpx<T> pa = cmk((T)(a.x - dr * kx), (T)(a.y - di * kx));
        const cpx<T> pb = cmk((T)(b.x - dr * ky), (T)(b.y - di * ky));
        const cpx<T> pc = cmk((T)(cc.x - dr * kzz), (T)(cc.y - di * kzz));
        Fh[0][i] = pa;
        Fh[1][i] = pb;
        Fh[2][i] = pc;
        if (ikx == 0 && iky == 0 && kz == 0) {
            dc[0] = pa.x;
            dc[1] = pb.x;
            dc[2] = pc.x;
        }
        double ore[3], oim[3];
        ore[0] = -(ky * (double)pc.y - kzz * (double)pb.y);
        oim[0] = (ky * (double)pc.x - kzz * (double)pb.x);
        ore[1] = -(kzz * (double)pa.y - kx * (double)pc.y);
        oim[1] = (kzz * (double)pa.x - kx * (double)pc.x);
        ore[2] = -(kx * (double)pb.y - ky * (double)pa.y);
        oim[2] = (kx * (double)pb.x - ky * (double)pa.x);
        int c0, c1;
        if (ikx != 0) {
            c0 = 1;
            c1 = 2;
        } else if (iky != 0) {
            c0 = 2;
            c1 = 0;
        } else {
            c0 = 0;
            c1 = 1;
        }
        a0[idx] = cmk((T)ore[c0], (T)oim[c0]);
        a1[idx] = cmk((T)ore[c1], (T)oim[c1]);
    }
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int p = threadIdx.x + i * NT;
            if (p < C * L) s[(p / C) * BS + p % C] = Fh[d][i];
        }
        __syncthreads();
        fft_cols_ip<T, L, C, BS, NT, false>(s, tw);
#pragma unroll
        for (int i = 0; i < NPT; i++) {
            const int e = threadIdx.x + i * NT;
            const int n = e / C, c = e - n * C;
            if (e < C * L && kz0 + c < g.trn_z) F[d][base + n * lstride + c] = s[n * BS + c];
        }
        __syncthreads();
    }
}

// zero zeta outside the kept modes (the reference's state there is
// round-off after U2zeta_k; here it is exactly zero)
template <typename T>
__global__ void k_zeta_clean(Grid g, cpx<T> *__restrict__ z0, cpx<T> *__restrict__ z1) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nspec) return;
    const int kz = (int)(idx % g.nzp);
    const long long rest = idx / g.nzp;
    const int j = (int)(rest % g.ny), i = (int)(rest / g.ny);
    if (kz < g.nzh && !kept_mode(g, kx_of(g, i), ky_of(g, j), kz)) {
        z0[idx] = cmk((T)0, (T)0);
        z1[idx] = cmk((T)0, (T)0);
    }
}

// zeta += (a0, a1) and uk_dc += dc: the projected slip field's vorticity
// and mean velocity, added after the slip iteration.
// With kept_only, (a0, a1) are read at the 2/3-kept modes only (they hold
// nothing else).
template <typename T>
__global__ void k_zeta_add(Grid g, cpx<T> *__restrict__ z0, cpx<T> *__restrict__ z1, const cpx<T> *__restrict__ a0,
                           const cpx<T> *__restrict__ a1, double *__restrict__ ukdc, const double *__restrict__ dc,
                           int kept_only) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0)
        for (int d = 0; d < 3; d++) ukdc[d] += dc[d];
    if (idx >= g.nspec || (int)(idx % g.nzp) >= g.nzh) return;
    if (kept_only) {
        const int       kz   = (int)(idx % g.nzp);
        const long long rest = idx / g.nzp;
        const int       j = (int)(rest % g.ny), i = (int)(rest / g.ny);
        if (!kept_mode(g, kx_of(g, i), ky_of(g, j), kz)) return;
    }
    z0[idx] = cadd(z0[idx], a0[idx]);
    z1[idx] = cadd(z1[idx], a1[idx]);
}

// ------------------------------------------------------------ launchers
#define KH_FOR_EACH_L(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

static inline int nky(const Grid &g) { return 2 * g.trn_y - 1; }

template <typename T, bool FWD>
hipError_t cols_y(const Grid &g, const Comps<T> &data, int ncomp, int in_mask, int out_mask, hipStream_t s,
                  const int *skip) {
    constexpr int C  = ColTile<T>::C;
    const int     nt = (g.trn_z + C - 1) / C;
    switch (g.ny) {
#define KH_CASE(n)                                                                                                 \
    case n:                                                                                                        \
        hipLaunchKernelGGL((k_cols<T, n, FWD, false>), dim3(g.nx * nt, ncomp, batch_z(g)), dim3(PassThreads<T>::value), 0, s, \
                           data, g, g.trn_z, in_mask, out_mask, skip);                                             \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T, bool FWD>
hipError_t cols_x(const Grid &g, const Comps<T> &data, int ncomp, int in_mask, int out_mask, hipStream_t s) {
    constexpr int C  = ColTile<T>::C;
    const int     nt = (g.trn_z + C - 1) / C;
    switch (g.nx) {
#define KH_CASE(n)                                                                                                  \
    case n:                                                                                                         \
        hipLaunchKernelGGL((k_cols<T, n, FWD, true>), dim3(nky(g) * nt, ncomp, batch_z(g)), dim3(PassThreads<T>::value), 0, s, \
                           data, g, g.trn_z, in_mask, out_mask, nullptr);                                           \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t x_zeta2u_inv(const Grid &g, const cpx<T> *z0, const cpx<T> *z1, cpx<T> *const u[3], double *ukdc,
                        int fixmask, hipStream_t s) {
    constexpr int C  = XZTile<T>::C;
    const int     nt = (g.trn_z + C - 1) / C;
    switch (g.nx) {
#define KH_CASE(n)                                                                                                   \
    case n:                                                                                                          \
        hipLaunchKernelGGL((k_x_zeta2u_inv<T, n>), dim3(nky(g) * nt, 1, batch_z(g)), dim3(XZTile<T>::NT), 0, s, g, z0, z1, u[0], u[1],    \
                           u[2], ukdc, fixmask);                                                                     \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t z_adv(const Grid &g, cpx<T> *const s3[3], cpx<T> *const p5[5], hipStream_t s) {
    const long long nrows = (long long)g.nx * g.ny;
    const T         scale = (T)(1.0 / ((double)g.nx * (double)g.ny * (double)g.nz));
    if (zw_supported(g.nz)) return zw_adv<T>(g, s3, p5, scale, s);
    switch (g.nz) {
#define KH_CASE(n)                                                                                                 \
    case n: {                                                                                                      \
        constexpr int P = (ZTile<T>::ELEM / n) > 0 ? ZTile<T>::ELEM / n : 1;                                       \
        hipLaunchKernelGGL((k_z_adv<T, n>), dim3((unsigned)((nrows + 2 * P - 1) / (2 * P)), 1, batch_z(g)), dim3(PassThreads<T>::value), 0, s, g, \
                           s3[0], s3[1], s3[2], p5[0], p5[1], p5[2], p5[3], p5[4], scale);                         \
        return hipGetLastError();                                                                                  \
    }
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

bool z_force_clears(const Grid &g) { return zw_supported(g.nz); }

template <typename T>
hipError_t z_force(const Grid &g, T *const up[3], T *phis, T *const us[3], cpx<T> *const f[3], int clear,
                   int *tick, hipStream_t s, T *const us2[3], const int *skip, Occ occ) {
    const long long nrows = (long long)g.nx * g.ny;
    if (zw_supported(g.nz)) return zw_force<T>(g, up, phis, us, f, clear, tick, s, us2, skip, occ);
    if (clear || tick || us2 || skip || occ.w) return hipErrorInvalidValue;
    T *const none[3] = {nullptr, nullptr, nullptr};
    if (!us) us = none;
    switch (g.nz) {
#define KH_CASE(n)                                                                                                   \
    case n: {                                                                                                        \
        constexpr int P = (ZTile<T>::ELEM / n) > 0 ? ZTile<T>::ELEM / n : 1;                                         \
        hipLaunchKernelGGL((k_z_force<T, n>), dim3((unsigned)((nrows + 2 * P - 1) / (2 * P)), 1, batch_z(g)), dim3(PassThreads<T>::value), 0, s, g, \
                           up[0], up[1], up[2], phis, us[0], us[1], us[2], f[0], f[1], f[2]);                        \
        return hipGetLastError();                                                                                    \
    }
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t x_adv_ns(const Grid &g, const Comps<T> &prod, cpx<T> *z0, cpx<T> *z1, double dmy0, double dmy1,
                    hipStream_t s) {
    constexpr int C  = XTile<T>::C;
    const int     nt = (g.trn_z + C - 1) / C;
    switch (g.nx) {
#define KH_CASE(n)                                                                                                   \
    case n:                                                                                                          \
        hipLaunchKernelGGL((k_x_adv_ns<T, n>), dim3(nky(g) * nt, 1, batch_z(g)), dim3(XTile<T>::NT), 0, s, g, prod, z0, z1, \
                           dmy0, dmy1);                                                                              \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t x_f_zeta(const Grid &g, cpx<T> *const f[3], cpx<T> *z0, cpx<T> *z1, double *ukdc, hipStream_t s) {
    constexpr int C  = XTile<T>::C;
    const int     nt = (g.trn_z + C - 1) / C;
    switch (g.nx) {
#define KH_CASE(n)                                                                                                 \
    case n:                                                                                                        \
        hipLaunchKernelGGL((k_x_f_zeta<T, n>), dim3(nky(g) * nt, 1, batch_z(g)), dim3(XTile<T>::NT), 0, s, g, f[0], f[1], f[2], z0, z1, \
                           ukdc);                                                                                  \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t x_slip(const Grid &g, cpx<T> *const f[3], cpx<T> *a0, cpx<T> *a1, double *dc, const int *skip,
                  hipStream_t s) {
    constexpr int C  = XTile<T>::C;
    const int     nt = (g.trn_z + C - 1) / C;
    switch (g.nx) {
#define KH_CASE(n)                                                                                                  \
    case n:                                                                                                         \
        hipLaunchKernelGGL((k_x_slip<T, n>), dim3(nky(g) * nt), dim3(XTile<T>::NT), 0, s, g, f[0], f[1], f[2], a0, a1, \
                           dc, skip);                                                                               \
        return hipGetLastError();
        KH_FOR_EACH_L(KH_CASE)
#undef KH_CASE
        default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t zeta_add(const Grid &g, cpx<T> *z0, cpx<T> *z1, const cpx<T> *a0, const cpx<T> *a1, double *ukdc,
                    const double *dc, hipStream_t s, int kept_only) {
    hipLaunchKernelGGL(k_zeta_add<T>, dim3((unsigned)((g.nspec + 255) / 256)), dim3(256), 0, s, g, z0, z1, a0, a1, ukdc,
                       dc, kept_only);
    return hipGetLastError();
}

template <typename T>
hipError_t zeta_clean(const Grid &g, cpx<T> *z0, cpx<T> *z1, hipStream_t s) {
    hipLaunchKernelGGL(k_zeta_clean<T>, dim3((unsigned)((g.nspec + 255) / 256)), dim3(256), 0, s, g, z0, z1);
    return hipGetLastError();
}

#define KH_INST(T)                                                                                                  \
    template hipError_t cols_y<T, true>(const Grid &, const Comps<T> &, int, int, int, hipStream_t, const int *);   \
    template hipError_t cols_y<T, false>(const Grid &, const Comps<T> &, int, int, int, hipStream_t, const int *);  \
    template hipError_t cols_x<T, true>(const Grid &, const Comps<T> &, int, int, int, hipStream_t);                \
    template hipError_t cols_x<T, false>(const Grid &, const Comps<T> &, int, int, int, hipStream_t);               \
    template hipError_t x_zeta2u_inv<T>(const Grid &, const cpx<T> *, const cpx<T> *, cpx<T> *const[3], double *,   \
                                        int, hipStream_t);                                                          \
    template hipError_t z_adv<T>(const Grid &, cpx<T> *const[3], cpx<T> *const[5], hipStream_t);                   \
    template hipError_t z_force<T>(const Grid &, T *const[3], T *, T *const[3], cpx<T> *const[3], int, int *,          \
                                   hipStream_t, T *const[3], const int *, Occ);                                    \
    template hipError_t x_slip<T>(const Grid &, cpx<T> *const[3], cpx<T> *, cpx<T> *, double *, const int *,          \
                                  hipStream_t);                                                                    \
    template hipError_t x_f_zeta<T>(const Grid &, cpx<T> *const[3], cpx<T> *, cpx<T> *, double *, hipStream_t);    \
    template hipError_t x_adv_ns<T>(const Grid &, const Comps<T> &, cpx<T> *, cpx<T> *, double, double, hipStream_t); \
    template hipError_t zeta_clean<T>(const Grid &, cpx<T> *, cpx<T> *, hipStream_t);                               \
    template hipError_t zeta_add<T>(const Grid &, cpx<T> *, cpx<T> *, const cpx<T> *, const cpx<T> *, double *,      \
                                    const double *, hipStream_t, int);
KH_INST(double)
KH_INST(float)
#undef KH_INST

}  // namespace kh
