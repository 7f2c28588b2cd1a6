// kh_spectral.hip -- pointwise spectral and real-space operators of the
// SPM fluid step, on the device layouts of kh_fft.hip. Wavenumbers are
// computed from the index on the fly (no K tables in HBM); the formulas and
// their operation order follow the reference lines cited per kernel.
#include "kh_launch.h"
#include "kh_modes.h"

namespace kh {

constexpr int PW_NT = 256;

// idx over nx*ny*nzp; kz >= nzh are the row padding (never touched)
__device__ __forceinline__ void spec_index(const Grid &g, long long idx, int &i, int &j, int &kz) {
    kz             = (int)(idx % g.nzp);
    long long rest = idx / g.nzp;
    j              = (int)(rest % g.ny);
    i              = (int)(rest / g.ny);
}

static inline unsigned nblocks(long long n) { return (unsigned)((n + PW_NT - 1) / PW_NT); }

// Zeta_k2u_k (fft_wrapper.cxx:591-688). With `truncate` set it also applies
// the 2/3 truncation of zeta that precedes it in the step
// (Truncate_two_third_rule, fft_wrapper.h:754-796, called at
// operate_omega.cxx:90-92 and sp_3d_ns.cxx:54-56): truncated modes give
// u_hat = 0. fixmask applies the FIX_CELL zeroing of uk_dc (sp_3d_ns.cxx:46-52).
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_zeta_to_u(Grid g, const cpx<T> *__restrict__ z0,
                                                     const cpx<T> *__restrict__ z1, cpx<T> *__restrict__ u0,
                                                     cpx<T> *__restrict__ u1, cpx<T> *__restrict__ u2,
                                                     double *__restrict__ ukdc, int fixmask, int truncate) {
    const long long idx = (long long)blockIdx.x * PW_NT + threadIdx.x;
    if (idx >= g.nspec) return;
    int i, j, kz;
    spec_index(g, idx, i, j, kz);
    if (kz >= g.nzh) return;
    const int ikx = kx_of(g, i), iky = ky_of(g, j);
    cpx<T>    r0 = cmk((T)0, (T)0), r1 = r0, r2 = r0;
    if (idx == 0) {
        double uk[3];
        for (int d = 0; d < 3; d++) {
            if (fixmask & (1 << d)) ukdc[d] = 0.0;
            uk[d] = ukdc[d];
        }
        r0 = cmk((T)uk[0], (T)0);
        r1 = cmk((T)uk[1], (T)0);
        r2 = cmk((T)uk[2], (T)0);
    } else if (!truncate || kept_mode(g, ikx, iky, kz)) {
        const double kx = g.wave_x * ikx, ky = g.wave_y * iky, kz_ = g.wave_z * kz;
        const double k2  = kx * kx + ky * ky + kz_ * kz_;
        const double ik2 = (k2 > 0.0) ? 1. / k2 : 0.0;
        const cpx<T> d1 = z0[idx], d2 = z1[idx];
        double o0r, o0i, o1r, o1i, o2r, o2i;
        if (ikx != 0) {
            o0r = -(1. / kx) * (ky * (double)d1.x + kz_ * (double)d2.x);
            o0i = -(1. / kx) * (ky * (double)d1.y + kz_ * (double)d2.y);
            o1r = d1.x;
            o1i = d1.y;
            o2r = d2.x;
            o2i = d2.y;
        } else if (iky != 0) {
            const double dmy = -(kz_ / ky);
            o0r              = d2.x;
            o0i              = d2.y;
            o1r              = dmy * (double)d1.x;
            o1i              = dmy * (double)d1.y;
            o2r              = d1.x;
            o2i              = d1.y;
        } else {
            o0r = d1.x;
            o0i = d1.y;
            o1r = d2.x;
            o1i = d2.y;
            o2r = 0.0;
            o2i = 0.0;
        }
        const double qx = kx * ik2, qy = ky * ik2, qz = kz_ * ik2;
        r0 = cmk((T)(-(qy * o2i - qz * o1i)), (T)(qy * o2r - qz * o1r));
        r1 = cmk((T)(-(qz * o0i - qx * o2i)), (T)(qz * o0r - qx * o2r));
        r2 = cmk((T)(-(qx * o1i - qy * o0i)), (T)(qx * o1r - qy * o0r));
    }
    u0[idx] = r0;
    u1[idx] = r1;
    u2[idx] = r2;
}

// Real-space quadratic products of U2advection_k (operate_omega.cxx:31-47).
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_products(long long n, const T *__restrict__ u0, const T *__restrict__ u1,
                                                    const T *__restrict__ u2, T *__restrict__ p0, T *__restrict__ p1,
                                                    T *__restrict__ p2, T *__restrict__ p3, T *__restrict__ p4) {
    const long long i = (long long)blockIdx.x * PW_NT + threadIdx.x;
    if (i >= n) return;
    const T a = u0[i], b = u1[i], c = u2[i];
    p0[i]     = b * c;
    p1[i]     = c * a;
    p2[i]     = a * b;
    p3[i]     = (b - c) * (b + c);
    p4[i]     = (c - a) * (c + a);
}

// Spectral part of U2advection_k (operate_omega.cxx:57-82), the reduction
// Omega_k2zeta_k (fft_wrapper.cxx:227-263) and the slaved-Euler update of
// NS_solver_slavedEuler (fluid_solver.cxx:355-373) in one pass over the
// kept modes; f_ns0 is never stored.
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_adv_ns(Grid g, const cpx<T> *__restrict__ P0, const cpx<T> *__restrict__ P1,
                                                  const cpx<T> *__restrict__ P2, const cpx<T> *__restrict__ P3,
                                                  const cpx<T> *__restrict__ P4, cpx<T> *__restrict__ z0,
                                                  cpx<T> *__restrict__ z1, double dmy0, double dmy1) {
    if (batch_skip(g)) return;
    P0 = brun(P0, g), P1 = brun(P1, g), P2 = brun(P2, g), P3 = brun(P3, g), P4 = brun(P4, g);
    z0 = brun(z0, g), z1 = brun(z1, g);
    // one thread per kept mode: (kx, ky) kept rows, kz < TRN_Z
    const long long t  = (long long)blockIdx.x * PW_NT + threadIdx.x;
    const int       ky = nkept(g.trn_y);
    if (t >= (long long)nkept(g.trn_x) * ky * g.trn_z) return;
    const int       kz   = (int)(t % g.trn_z);
    const long long rest = t / g.trn_z;
    const int       j    = kept_index((int)(rest % ky), g.ny, g.trn_y), i = kept_index((int)(rest / ky), g.nx, g.trn_x);
    const int       ikx = kx_of(g, i), iky = ky_of(g, j);
    const long long idx = ((long long)i * g.ny + j) * g.nzp + kz;
    const cpx<T>    a[5] = {P0[idx], P1[idx], P2[idx], P3[idx], P4[idx]};
    slaved_euler_mode<T>(g, ikx, iky, kz, a, dmy0, dmy1, z0[idx], z1[idx]);
}

// Truncation + Solenoidal_uk (operate_omega.h:105-112, operate_omega.cxx:171-195).
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_solenoidal(Grid g, cpx<T> *__restrict__ u0, cpx<T> *__restrict__ u1,
                                                      cpx<T> *__restrict__ u2) {
    const long long idx = (long long)blockIdx.x * PW_NT + threadIdx.x;
    if (idx >= g.nspec) return;
    int i, j, kz;
    spec_index(g, idx, i, j, kz);
    if (kz >= g.nzh) return;
    const int ikx = kx_of(g, i), iky = ky_of(g, j);
    if (!kept_mode(g, ikx, iky, kz)) {
        u0[idx] = u1[idx] = u2[idx] = cmk((T)0, (T)0);
        return;
    }
    const double kx = ikx * g.wave_x, ky = iky * g.wave_y, kzz = kz * g.wave_z;
    const double K2  = (g.wave_x * ikx) * (g.wave_x * ikx) + (g.wave_y * iky) * (g.wave_y * iky) +
                      (g.wave_z * kz) * (g.wave_z * kz);
    const double IK2 = (K2 > 0.0) ? 1. / K2 : 0.0;
    cpx<T>       a = u0[idx], b = u1[idx], c = u2[idx];
    const double dr = IK2 * ((double)a.x * kx + (double)b.x * ky + (double)c.x * kzz);
    const double di = IK2 * ((double)a.y * kx + (double)b.y * ky + (double)c.y * kzz);
    u0[idx]         = cmk((T)(a.x - dr * kx), (T)(a.y - di * kx));
    u1[idx]         = cmk((T)(b.x - dr * ky), (T)(b.y - di * ky));
    u2[idx]         = cmk((T)(c.x - dr * kzz), (T)(c.y - di * kzz));
}

// U_k2zeta_k (fft_wrapper.cxx:407-476): zeta from i k x u_hat; uk_dc from u_hat(0).
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_u_to_zeta(Grid g, const cpx<T> *__restrict__ u0,
                                                     const cpx<T> *__restrict__ u1, const cpx<T> *__restrict__ u2,
                                                     cpx<T> *__restrict__ z0, cpx<T> *__restrict__ z1,
                                                     double *__restrict__ ukdc) {
    const long long idx = (long long)blockIdx.x * PW_NT + threadIdx.x;
    if (idx >= g.nspec) return;
    int i, j, kz;
    spec_index(g, idx, i, j, kz);
    if (kz >= g.nzh) return;
    const int    ikx = kx_of(g, i), iky = ky_of(g, j);
    const double ks0 = ikx * g.wave_x, ks1 = iky * g.wave_y, ks2 = kz * g.wave_z;
    const cpx<T> a = u0[idx], b = u1[idx], c = u2[idx];
    if (idx == 0) {
        ukdc[0] = a.x;
        ukdc[1] = b.x;
        ukdc[2] = c.x;
    }
    double ore[3], oim[3];
    ore[0] = -(ks1 * (double)c.y - ks2 * (double)b.y);
    oim[0] = (ks1 * (double)c.x - ks2 * (double)b.x);
    ore[1] = -(ks2 * (double)a.y - ks0 * (double)c.y);
    oim[1] = (ks2 * (double)a.x - ks0 * (double)c.x);
    ore[2] = -(ks0 * (double)b.y - ks1 * (double)a.y);
    oim[2] = (ks0 * (double)b.x - ks1 * (double)a.x);
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

// Make_f_particle_dt_nonsole (f_particle.h:35-54), phi = min(phi_sum, 1)
// (make_phi.cxx:248-259) evaluated on the fly.
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_f_nonsole(long long n, const T *__restrict__ up0, const T *__restrict__ up1,
                                                     const T *__restrict__ up2, const T *__restrict__ phisum,
                                                     const T *__restrict__ u0, const T *__restrict__ u1,
                                                     const T *__restrict__ u2, T *__restrict__ f0, T *__restrict__ f1,
                                                     T *__restrict__ f2) {
    const long long i = (long long)blockIdx.x * PW_NT + threadIdx.x;
    if (i >= n) return;
    const T ph = fmin(phisum[i], (T)1);
    f0[i]      = up0[i] - ph * u0[i];
    f1[i]      = up1[i] - ph * u1[i];
    f2[i]      = up2[i] - ph * u2[i];
}

// Add_f_particle (f_particle.h:68-84): u += f.
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_add3(long long n, T *__restrict__ u0, T *__restrict__ u1,
                                                T *__restrict__ u2, const T *__restrict__ f0,
                                                const T *__restrict__ f1, const T *__restrict__ f2) {
    const long long i = (long long)blockIdx.x * PW_NT + threadIdx.x;
    if (i >= n) return;
    u0[i] += f0[i];
    u1[i] += f1[i];
    u2[i] += f2[i];
}

// 16-byte stores; fields are hipMalloc'd (256-byte aligned)
template <typename T>
__global__ __launch_bounds__(PW_NT) void k_fill(long long n, T *__restrict__ p, T v, Grid bg) {
    if (batch_skip(bg)) return;
    p                  = brun(p, bg);
    constexpr int   W  = 16 / sizeof(T);
    const long long i  = ((long long)blockIdx.x * PW_NT + threadIdx.x) * W;
    if (i + W <= n) {
        T tmp[W];
#pragma unroll
        for (int k = 0; k < W; k++) tmp[k] = v;
        *reinterpret_cast<uint4 *>(p + i) = *reinterpret_cast<const uint4 *>(tmp);
    } else {
        for (long long k = i; k < n; k++) p[k] = v;
    }
}

// ------------------------------------------------------------ launchers
template <typename T>
hipError_t zeta_to_u(const Grid &g, const cpx<T> *z0, const cpx<T> *z1, cpx<T> *u0, cpx<T> *u1, cpx<T> *u2,
                     double *ukdc, int fixmask, int truncate, hipStream_t s) {
    hipLaunchKernelGGL(k_zeta_to_u<T>, dim3(nblocks(g.nspec)), dim3(PW_NT), 0, s, g, z0, z1, u0, u1, u2, ukdc, fixmask,
                       truncate);
    return hipGetLastError();
}
template <typename T>
hipError_t products(long long n, const T *u0, const T *u1, const T *u2, T *const p[5], hipStream_t s) {
    hipLaunchKernelGGL(k_products<T>, dim3(nblocks(n)), dim3(PW_NT), 0, s, n, u0, u1, u2, p[0], p[1], p[2], p[3],
                       p[4]);
    return hipGetLastError();
}
template <typename T>
hipError_t adv_ns(const Grid &g, cpx<T> *const P[5], cpx<T> *z0, cpx<T> *z1, double dmy0, double dmy1,
                  hipStream_t s) {
    const long long nk = (long long)(2 * g.trn_x - 1) * (2 * g.trn_y - 1) * g.trn_z;
    hipLaunchKernelGGL(k_adv_ns<T>, dim3(nblocks(nk), 1, batch_z(g)), dim3(PW_NT), 0, s, g, P[0], P[1], P[2], P[3], P[4],
                       z0, z1, dmy0, dmy1);
    return hipGetLastError();
}
template <typename T>
hipError_t solenoidal(const Grid &g, cpx<T> *u0, cpx<T> *u1, cpx<T> *u2, hipStream_t s) {
    hipLaunchKernelGGL(k_solenoidal<T>, dim3(nblocks(g.nspec)), dim3(PW_NT), 0, s, g, u0, u1, u2);
    return hipGetLastError();
}
template <typename T>
hipError_t u_to_zeta(const Grid &g, const cpx<T> *u0, const cpx<T> *u1, const cpx<T> *u2, cpx<T> *z0, cpx<T> *z1,
                     double *ukdc, hipStream_t s) {
    hipLaunchKernelGGL(k_u_to_zeta<T>, dim3(nblocks(g.nspec)), dim3(PW_NT), 0, s, g, u0, u1, u2, z0, z1, ukdc);
    return hipGetLastError();
}
template <typename T>
hipError_t f_nonsole(long long n, T *const up[3], const T *phisum, T *const u[3], T *const f[3], hipStream_t s) {
    hipLaunchKernelGGL(k_f_nonsole<T>, dim3(nblocks(n)), dim3(PW_NT), 0, s, n, up[0], up[1], up[2], phisum, u[0], u[1],
                       u[2], f[0], f[1], f[2]);
    return hipGetLastError();
}
template <typename T>
hipError_t add3(long long n, T *const u[3], T *const f[3], hipStream_t s) {
    hipLaunchKernelGGL(k_add3<T>, dim3(nblocks(n)), dim3(PW_NT), 0, s, n, u[0], u[1], u[2], f[0], f[1], f[2]);
    return hipGetLastError();
}
template <typename T>
hipError_t fill(long long n, T *p, T v, hipStream_t s, const Grid *batch) {
    constexpr int W  = 16 / sizeof(T);
    const Grid    bg = batch ? *batch : Grid{};
    hipLaunchKernelGGL(k_fill<T>, dim3(nblocks((n + W - 1) / W), 1, batch_z(bg)), dim3(PW_NT), 0, s, n, p, v, bg);
    return hipGetLastError();
}

#define KH_INST(T)                                                                                                 \
    template hipError_t zeta_to_u<T>(const Grid &, const cpx<T> *, const cpx<T> *, cpx<T> *, cpx<T> *, cpx<T> *,   \
                                     double *, int, int, hipStream_t);                                             \
    template hipError_t products<T>(long long, const T *, const T *, const T *, T *const[5], hipStream_t);         \
    template hipError_t adv_ns<T>(const Grid &, cpx<T> *const[5], cpx<T> *, cpx<T> *, double, double, hipStream_t); \
    template hipError_t solenoidal<T>(const Grid &, cpx<T> *, cpx<T> *, cpx<T> *, hipStream_t);                    \
    template hipError_t u_to_zeta<T>(const Grid &, const cpx<T> *, const cpx<T> *, const cpx<T> *, cpx<T> *,       \
                                     cpx<T> *, double *, hipStream_t);                                             \
    template hipError_t f_nonsole<T>(long long, T *const[3], const T *, T *const[3], T *const[3], hipStream_t);    \
    template hipError_t add3<T>(long long, T *const[3], T *const[3], hipStream_t);                                 \
    template hipError_t fill<T>(long long, T *, T, hipStream_t, const Grid *);
KH_INST(double)
KH_INST(float)
#undef KH_INST

}  // namespace kh
