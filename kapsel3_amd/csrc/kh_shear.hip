// kh_shear.hip -- the Lees-Edwards sheared step's fluid kernels
// (SW_EQ == Shear_Navier_Stokes_Lees_Edwards, Time_evolution_hydro_OBL,
// sp_3d_ns.cxx:328-433).
//
// The fluid lives on an oblique grid that follows the mean shear flow
// rate*(y - LY/2): velocities are contravariant components on that grid and
// g (degree_oblique) is the strain, wrapped into (-NX/NY, NX/NY) by a remap.
// Wavenumbers are the covariant integers of the grid; the squared modulus is
// K2 = kx^2 + (ky - g kx)^2 + kz^2 (Update_K2_OBL, fluid_solver.h:342-360),
// computed on the fly from g like every other k-dependent factor here.
// Spectral arithmetic is FP64 on T-precision storage, as in kh_spectral.hip.
#include "kh_launch.h"

namespace kh {

constexpr int SH_NT = 256;

static inline unsigned sh_blocks(long long n) { return (unsigned)((n + SH_NT - 1) / SH_NT); }

// Grid axes are at most 512 points, so every element index fits 32 bits:
// index arithmetic is unsigned 32-bit (a 64-bit division is a long
// instruction sequence on the GPU).
__device__ __forceinline__ void sh_spec_index(const Grid &g, unsigned idx, int &i, int &j, int &kz) {
    kz                  = (int)(idx % (unsigned)g.nzp);
    const unsigned rest = idx / (unsigned)g.nzp;
    j                   = (int)(rest % (unsigned)g.ny);
    i                   = (int)(rest / (unsigned)g.ny);
}

// t-th kept mode (|kx| < trn_x, |ky| < trn_y, kz < trn_z), kz fastest: a
// launch over sh_nkept(g) threads visits each kept mode once.
__device__ __forceinline__ void sh_kept_index(const Grid &g, unsigned t, int &i, int &j, int &kz) {
    const unsigned nky  = (unsigned)nkept(g.trn_y);
    kz                  = (int)(t % (unsigned)g.trn_z);
    const unsigned rest = t / (unsigned)g.trn_z;
    j                   = kept_index((int)(rest % nky), g.ny, g.trn_y);
    i                   = kept_index((int)(rest / nky), g.nx, g.trn_x);
}
static inline long long sh_nkept(const Grid &g) { return (long long)(2 * g.trn_x - 1) * (2 * g.trn_y - 1) * g.trn_z; }

// Real-space kernels walk whole rows (fixed i, j; k contiguous), SH_RPW rows
// per workgroup, so a row's indices and interpolation weights are worked out
// once per row.
constexpr int SH_RPW = 8;
static inline unsigned sh_row_blocks(const Grid &g) {
    return (unsigned)(((long long)g.nx * g.ny + SH_RPW - 1) / SH_RPW);
}

__device__ __forceinline__ double k2_obl(const Grid &g, int ikx, int iky, int kz, double gd) {
    const double a = g.wave_x * ikx, b = g.wave_y * iky - g.wave_x * ikx * gd, c = g.wave_z * kz;
    return a * a + b * b + c * c;
}

// contra2co_single (fft_wrapper.h:611-621)
__device__ __forceinline__ void contra2co(double gd, double *v) {
    const double a = v[0], b = v[1];
    v[0]           = a + gd * b;
    v[1]           = gd * a + (1. + gd * gd) * b;
}

// Zeta_k2u_k_OBL (fft_wrapper.cxx:690-768): contravariant u_hat from zeta,
// u_hat(0) = uk_dc. With `truncate` the 2/3 rule is applied to zeta first,
// in place (operate_omega.cxx:103-105); w1 (optional) receives a copy of
// u_hat^y (Zeta_k2u_cpuky, operate_omega.h:153-157). fixmask zeroes uk_dc
// components first (FIX_CELL, sp_3d_ns.cxx:339-345). Only the kept modes of
// u_hat and w1 are written: their consumers (the pruned inverse transform,
// the slaved-Euler pass) read no others, and zeta is zero outside them.
// truncate < 0: every mode is converted (zeta as given). With truncate == 0
// the launch covers the kept modes only (sh_kept_index).
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_zeta_to_u(Grid g, cpx<T> *__restrict__ z0, cpx<T> *__restrict__ z1,
                                                         cpx<T> *__restrict__ u0, cpx<T> *__restrict__ u1,
                                                         cpx<T> *__restrict__ u2, cpx<T> *__restrict__ w1,
                                                         double *__restrict__ ukdc, int fixmask, int truncate,
                                                         double gd, unsigned nthreads) {
    const unsigned t = blockIdx.x * SH_NT + threadIdx.x;
    if (t >= nthreads) return;
    int i, j, kz;
    if (truncate == 0) {
        sh_kept_index(g, t, i, j, kz);
    } else {
        sh_spec_index(g, t, i, j, kz);
        if (kz >= g.nzh) return;
    }
    const unsigned idx = ((unsigned)i * g.ny + j) * g.nzp + kz;
    const int      ikx = kx_of(g, i), iky = ky_of(g, j);
    cpx<T>    r0 = cmk((T)0, (T)0), r1 = r0, r2 = r0;
    if (truncate > 0 && !kept_mode(g, ikx, iky, kz)) {
        if (truncate) z0[idx] = z1[idx] = r0;
        return;
    }
    if (idx != 0) {
        const double kx = g.wave_x * ikx, ky = g.wave_y * iky, kzz = g.wave_z * kz;
        const double K2 = k2_obl(g, ikx, iky, kz, gd), ik2 = K2 > 0.0 ? 1.0 / K2 : 0.0;
        const cpx<T> d1 = z0[idx], d2 = z1[idx];
        double       ore[3], oim[3];
        if (ikx != 0) {
            ore[0] = -(1. / kx) * (ky * (double)d1.x + kzz * (double)d2.x);
            oim[0] = -(1. / kx) * (ky * (double)d1.y + kzz * (double)d2.y);
            ore[1] = d1.x;
            oim[1] = d1.y;
            ore[2] = d2.x;
            oim[2] = d2.y;
        } else if (kz != 0) {
            const double dmy = -(ky / kzz);
            ore[0]           = d2.x;
            oim[0]           = d2.y;
            ore[1]           = d1.x;
            oim[1]           = d1.y;
            ore[2]           = dmy * (double)d1.x;
            oim[2]           = dmy * (double)d1.y;
        } else {
            ore[0] = d1.x;
            oim[0] = d1.y;
            ore[1] = 0.0;
            oim[1] = 0.0;
            ore[2] = d2.x;
            oim[2] = d2.y;
        }
        contra2co(gd, ore);
        contra2co(gd, oim);
        const double qx = kx * ik2, qy = ky * ik2, qz = kzz * ik2;
        r0 = cmk((T)(-(qy * oim[2] - qz * oim[1])), (T)(qy * ore[2] - qz * ore[1]));
        r1 = cmk((T)(-(qz * oim[0] - qx * oim[2])), (T)(qz * ore[0] - qx * ore[2]));
        r2 = cmk((T)(-(qx * oim[1] - qy * oim[0])), (T)(qx * ore[1] - qy * ore[0]));
    }
    if (idx == 0) {
        double uk[3];
        for (int d = 0; d < 3; d++) {
            if (fixmask & (1 << d)) ukdc[d] = 0.0;
            uk[d] = ukdc[d];
        }
        r0 = cmk((T)uk[0], (T)0);
        r1 = cmk((T)uk[1], (T)0);
        r2 = cmk((T)uk[2], (T)0);
    }
    u0[idx] = r0;
    u1[idx] = r1;
    u2[idx] = r2;
    if (w1) w1[idx] = r1;
}

// Zeta_k2omega_k_OBL (fft_wrapper.cxx:820-865): contravariant omega_hat
// (real coefficients on each of re and im; the real part of mode 0
// zeroed), at the kept modes (zeta is truncated; the pruned inverse
// transform reads no others)
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_zeta_to_omega(Grid g, const cpx<T> *__restrict__ z0,
                                                             const cpx<T> *__restrict__ z1, cpx<T> *__restrict__ o0,
                                                             cpx<T> *__restrict__ o1, cpx<T> *__restrict__ o2,
                                                             unsigned nthreads) {
    const unsigned t = blockIdx.x * SH_NT + threadIdx.x;
    if (t >= nthreads) return;
    int i, j, kz;
    sh_kept_index(g, t, i, j, kz);
    const unsigned idx = ((unsigned)i * g.ny + j) * g.nzp + kz;
    const int      ikx = kx_of(g, i), iky = ky_of(g, j);
    const double kx = g.wave_x * ikx, ky = g.wave_y * iky, kzz = g.wave_z * kz;
    const cpx<T> d1 = z0[idx], d2 = z1[idx];
    cpx<T>       a, b, c;
    if (ikx != 0) {
        a = cmk((T)(-(1. / kx) * (ky * (double)d1.x + kzz * (double)d2.x)),
                (T)(-(1. / kx) * (ky * (double)d1.y + kzz * (double)d2.y)));
        b = d1;
        c = d2;
    } else if (kz != 0) {
        a = d2;
        b = d1;
        c = cmk((T)(-(ky / kzz) * (double)d1.x), (T)(-(ky / kzz) * (double)d1.y));
    } else {
        a = d1;
        b = cmk((T)0, (T)0);
        c = d2;
    }
    if (idx == 0) {
        a.x = (T)0;
        b.x = (T)0;
        c.x = (T)0;
    }
    o0[idx] = a;
    o1[idx] = b;
    o2[idx] = c;
}

// u x omega of contravariant u and omega: covariant components
// (operate_omega.cxx:113-137), in place into w.
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_cross(long long n, const T *__restrict__ u0, const T *__restrict__ u1,
                                                     const T *__restrict__ u2, T *__restrict__ w0,
                                                     T *__restrict__ w1, T *__restrict__ w2) {
    const long long i = (long long)blockIdx.x * SH_NT + threadIdx.x;
    if (i >= n) return;
    const T a = u0[i], b = u1[i], c = u2[i], x = w0[i], y = w1[i], z = w2[i];
    w0[i]     = b * z - c * y;
    w1[i]     = c * x - a * z;
    w2[i]     = a * y - b * x;
}

// The spectral end of Zeta_k2advection_k_OBL (operate_omega.cxx:139-154):
// the shear terms -2 rate u_hat^y (1, g, 0), the curl (U_k2rotation_k,
// fft_wrapper.cxx:893-935) and Omega_k2zeta_k_OBL (fft_wrapper.cxx:265-289),
// then the slaved-Euler update of NS_solver_slavedEuler_Shear_OBL
// (fluid_solver.cxx:411-437), on the kept modes only.
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_adv_ns(Grid g, const cpx<T> *__restrict__ F0,
                                                      const cpx<T> *__restrict__ F1, const cpx<T> *__restrict__ F2,
                                                      const cpx<T> *__restrict__ w1, cpx<T> *__restrict__ z0,
                                                      cpx<T> *__restrict__ z1, double rate, double gd, double dmy0,
                                                      double dmy1, unsigned nthreads) {
    const unsigned t = blockIdx.x * SH_NT + threadIdx.x;
    if (t >= nthreads) return;
    int i, j, kz;
    sh_kept_index(g, t, i, j, kz);
    const int      ikx = kx_of(g, i), iky = ky_of(g, j);
    const unsigned idx = ((unsigned)i * g.ny + j) * g.nzp + kz;
    const cpx<T>    w   = w1[idx];
    const double    s2  = 2. * rate;
    // F_hat (u x omega) less the shear terms, re/im
    double fr[3], fi[3];
    const cpx<T> a = F0[idx], b = F1[idx], c = F2[idx];
    fr[0] = (double)a.x - s2 * (double)w.x;
    fi[0] = (double)a.y - s2 * (double)w.y;
    fr[1] = (double)b.x - s2 * gd * (double)w.x;
    fi[1] = (double)b.y - s2 * gd * (double)w.y;
    fr[2] = c.x;
    fi[2] = c.y;
    const double k0 = ikx * g.wave_x, k1 = iky * g.wave_y, k2 = kz * g.wave_z;
    double       rr[3], ri[3];
    rr[0] = -(k1 * fi[2] - k2 * fi[1]);
    ri[0] = (k1 * fr[2] - k2 * fr[1]);
    rr[1] = -(k2 * fi[0] - k0 * fi[2]);
    ri[1] = (k2 * fr[0] - k0 * fr[2]);
    rr[2] = -(k0 * fi[1] - k1 * fi[0]);
    ri[2] = (k0 * fr[1] - k1 * fr[0]);
    int c0, c1;
    if (ikx != 0) {
        c0 = 1;
        c1 = 2;
    } else if (kz != 0) {
        c0 = 1;
        c1 = 0;
    } else {
        c0 = 0;
        c1 = 2;
    }
    const double K2  = k2_obl(g, ikx, iky, kz, gd);
    const double IK2 = K2 > 0.0 ? 1.0 / K2 : 0.0;
    const double dmy = exp(dmy0 * K2) - 1.;
    const cpx<T> x0 = z0[idx], x1 = z1[idx];
    z0[idx] = cmk((T)((double)x0.x + dmy * ((double)x0.x - dmy1 * IK2 * rr[c0])),
                  (T)((double)x0.y + dmy * ((double)x0.y - dmy1 * IK2 * ri[c0])));
    z1[idx] = cmk((T)((double)x1.x + dmy * ((double)x1.x - dmy1 * IK2 * rr[c1])),
                  (T)((double)x1.y + dmy * ((double)x1.y - dmy1 * IK2 * ri[c1])));
}

// Reset_U_OBL (fluid_solver.h:312-340): the field at strain +-NX/NY
// re-expressed on the grid at strain 0; flg 0 from +NX/NY, 1 from -NX/NY.
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_remap(Grid g, const T *__restrict__ u0, const T *__restrict__ u1,
                                                     const T *__restrict__ u2, T *__restrict__ o0, T *__restrict__ o1,
                                                     T *__restrict__ o2, double delta) {
    const long long t = (long long)blockIdx.x * SH_NT + threadIdx.x;
    if (t >= g.nreal) return;
    const int       k    = (int)(t % g.nz);
    const long long rest = t / g.nz;
    const int       j = (int)(rest % g.ny), i = (int)(rest / g.ny);
    double          sign = j - g.ny / 2;
    if (!(sign == 0)) sign = sign / fabs(sign);
    int io = (int)((int)(sign * (j - g.ny / 2)) * sign);
    io     = (int)fmod(i + delta * io + 4. * g.nx, (double)g.nx);
    const long long q = ((long long)io * g.ny + j) * g.nz + k;
    o0[q]             = (T)((double)u0[t] + delta * (double)u1[t]);
    o1[q]             = u1[t];
    o2[q]             = u2[t];
}

// obl2orth (fft_wrapper.h:276-288) / orth2obl (262-274)
__device__ __forceinline__ void obl2orth(const Grid &g, double gd, int j, int i, int &i_plus, int &i_oblique,
                                         double &alpha, double &beta) {
    const int delta_j = j - g.ny / 2;
    double    sign    = (gd >= 0.0 ? (double)delta_j : -(double)delta_j);
    if (!(delta_j == 0)) sign = sign / fabs(sign);
    const int io = (int)((int)(sign * gd * delta_j) * sign + sign);
    alpha        = (io - gd * delta_j) * sign;
    beta         = 1.0 - alpha;
    // (the indices are integers: fmod of their sums is the integer remainder)
    i_oblique    = (i + io + 4 * g.nx) % g.nx;
    i_plus       = (i + (int)sign + 2 * g.nx) % g.nx;
}
__device__ __forceinline__ void orth2obl(const Grid &g, double gd, int j, int i, int &i_oblique,
                                         int &i_oblique_plus, double &alpha, double &beta) {
    const int delta_j = j - g.ny / 2;
    double    sign    = (gd >= 0.0 ? (double)delta_j : -(double)delta_j);
    if (!(delta_j == 0)) sign = sign / fabs(sign);
    const int io   = (int)((int)(sign * gd * delta_j) * sign);
    alpha          = (gd * delta_j - io) * sign;
    beta           = 1.0 - alpha;
    i_oblique      = (i + io + 4 * g.nx) % g.nx;
    i_oblique_plus = (i_oblique + (int)sign + 4 * g.nx) % g.nx;
}

// U_oblique2u (fft_wrapper.h:480-505): oblique grid, contravariant
// components -> Cartesian grid and components, plus the mean flow. With a
// mask (phi_sum) only the points where it is nonzero are written: the
// step reads this field nowhere else (F_H gathers and f = u_p - phi u
// are zero-weighted there). With `acc`, every point's term of
// Calc_Reynolds_shear_stress (fluid_solver.h:165-183),
// (u_x - rate (j - NY/2) DX) u_y, is summed per workgroup into
// acc[blockIdx.x] (one partial per workgroup, summed in a fixed order by
// the record kernel).
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_to_cart(Grid g, const T *__restrict__ u0, const T *__restrict__ u1,
                                                       const T *__restrict__ u2, T *__restrict__ c0,
                                                       T *__restrict__ c1, T *__restrict__ c2, double gd,
                                                       double rate, const T *__restrict__ mask,
                                                       double *__restrict__ acc, double dx) {
    const unsigned nrow = (unsigned)g.nx * g.ny;
    double         rey  = 0.0;
    for (int w = 0; w < SH_RPW; w++) {
        const unsigned r = blockIdx.x * SH_RPW + w;
        if (r >= nrow) break;  // the same row for every thread
        const int i = (int)(r / (unsigned)g.ny), j = (int)(r - (unsigned)i * g.ny);
        int       ip, io;
        double    alpha, beta;
        obl2orth(g, gd, j, i, ip, io, alpha, beta);
        const unsigned rt = r * g.nz, rp = ((unsigned)ip * g.ny + j) * g.nz, ro = ((unsigned)io * g.ny + j) * g.nz;
        for (int k = threadIdx.x; k < g.nz; k += SH_NT) {
            const unsigned t = rt + k, qp = rp + k, qo = ro + k;
            const bool     put = !mask || mask[qo] != (T)0;
            if (!put && !acc) continue;
            const double a0 = beta * (double)u0[t] + alpha * (double)u0[qp];
            const double a1 = beta * (double)u1[t] + alpha * (double)u1[qp];
            // the reference adds the two terms to the stored (T) value in turn
            T       x = (T)a0;
            const T y = (T)a1;
            x         = (T)((double)x + gd * (double)y);
            x         = (T)((double)x + rate * (j - g.ny / 2.0));
            if (acc) rey += ((double)x - rate * ((double)(j - g.ny / 2) * dx)) * (double)y;
            if (put) {
                c0[qo] = x;
                c1[qo] = y;
                c2[qo] = (T)(beta * (double)u2[t] + alpha * (double)u2[qp]);
            }
        }
    }
    if (acc) {
        __shared__ double red[SH_NT / 64];
        for (int off = 32; off > 0; off >>= 1) rey += __shfl_down(rey, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rey;
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = 0.0;
            for (int w = 0; w < SH_NT / 64; w++) b += red[w];
            acc[blockIdx.x] = b;
        }
    }
}

// workgroups of k_obl_to_cart: the length of its `acc` array
long long obl_to_cart_blocks(const Grid &g) { return (long long)sh_row_blocks(g); }

// Make_f_particle_dt_nonsole (f_particle.h:35-54) on the Cartesian grid,
// U2u_oblique (fft_wrapper.h:448-478) onto the oblique grid and
// Add_f_particle (f_particle.h:68-84), per oblique point. Where phi_sum is
// zero u_p is zero too, so f = 0 there: such points are not read (the
// Cartesian u* may be unwritten at them), and an oblique point whose two
// source points both have f = 0 keeps its u (u + 0).
template <typename T>
__device__ __forceinline__ void f_at(const T *up0, const T *up1, const T *up2, const T *c0, const T *c1, const T *c2,
                                     T ps, unsigned q, T *f) {
    if (ps == (T)0) {
        f[0] = f[1] = f[2] = (T)0;
        return;
    }
    const T ph = fmin(ps, (T)1);
    f[0]       = up0[q] - ph * c0[q];
    f[1]       = up1[q] - ph * c1[q];
    f[2]       = up2[q] - ph * c2[q];
}
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_add_f(Grid g, const T *__restrict__ up0, const T *__restrict__ up1,
                                                     const T *__restrict__ up2, const T *__restrict__ phisum,
                                                     const T *__restrict__ c0, const T *__restrict__ c1,
                                                     const T *__restrict__ c2, T *__restrict__ u0,
                                                     T *__restrict__ u1, T *__restrict__ u2, double gd) {
    const unsigned nrow = (unsigned)g.nx * g.ny;
    for (int w = 0; w < SH_RPW; w++) {
        const unsigned r = blockIdx.x * SH_RPW + w;
        if (r >= nrow) return;  // the same row for every thread
        const int i = (int)(r / (unsigned)g.ny), j = (int)(r - (unsigned)i * g.ny);
        int       io, iop;
        double    alpha, beta;
        orth2obl(g, gd, j, i, io, iop, alpha, beta);
        const unsigned rt = r * g.nz, ra = ((unsigned)io * g.ny + j) * g.nz, rb = ((unsigned)iop * g.ny + j) * g.nz;
        for (int k = threadIdx.x; k < g.nz; k += SH_NT) {
            const unsigned t = rt + k, qa = ra + k, qb = rb + k;
            const T        pa = phisum[qa], pb = phisum[qb];
            if (pa == (T)0 && pb == (T)0) continue;
            T fa[3], fb[3];
            f_at(up0, up1, up2, c0, c1, c2, pa, qa, fa);
            f_at(up0, up1, up2, c0, c1, c2, pb, qb, fb);
            T       f0 = (T)(beta * (double)fa[0] + alpha * (double)fb[0]);
            const T f1 = (T)(beta * (double)fa[1] + alpha * (double)fb[1]);
            const T f2 = (T)(beta * (double)fa[2] + alpha * (double)fb[2]);
            f0         = (T)((double)f0 - gd * (double)f1);
            u0[t] += f0;
            u1[t] += f1;
            u2[t] += f2;
        }
    }
}

// Solenoidal_uk_OBL (operate_omega.cxx:197-227) and U_k2zeta_k_OBL
// (fft_wrapper.cxx:526-589) per mode; uk_dc from the projected u_hat(0).
// With `kept` the launch covers the kept modes only (sh_kept_index).
template <typename T>
__global__ __launch_bounds__(SH_NT) void k_obl_u_to_zeta(Grid g, const cpx<T> *__restrict__ u0,
                                                         const cpx<T> *__restrict__ u1, const cpx<T> *__restrict__ u2,
                                                         cpx<T> *__restrict__ z0, cpx<T> *__restrict__ z1,
                                                         double *__restrict__ ukdc, double gd, int kept,
                                                         unsigned nthreads) {
    const unsigned t = blockIdx.x * SH_NT + threadIdx.x;
    if (t >= nthreads) return;
    int i, j, kz;
    if (kept) {
        sh_kept_index(g, t, i, j, kz);
    } else {
        sh_spec_index(g, t, i, j, kz);
        if (kz >= g.nzh) return;
    }
    const unsigned idx = ((unsigned)i * g.ny + j) * g.nzp + kz;
    const int    ikx = kx_of(g, i), iky = ky_of(g, j);
    const double kx = ikx * g.wave_x, ky = iky * g.wave_y, kzz = kz * g.wave_z;
    const double kxc = (1. + gd * gd) * ikx * g.wave_x - gd * iky * g.wave_y;
    const double kyc = -gd * ikx * g.wave_x + iky * g.wave_y;
    const double K2 = k2_obl(g, ikx, iky, kz, gd), IK2 = K2 > 0.0 ? 1.0 / K2 : 0.0;
    const cpx<T> a = u0[idx], b = u1[idx], c = u2[idx];
    const double dr = IK2 * ((double)a.x * kx + (double)b.x * ky + (double)c.x * kzz);
    const double di = IK2 * ((double)a.y * kx + (double)b.y * ky + (double)c.y * kzz);
    // the projected field is stored in T before the curl reads it
    double ur[3] = {(double)(T)((double)a.x - dr * kxc), (double)(T)((double)b.x - dr * kyc),
                    (double)(T)((double)c.x - dr * kzz)};
    double ui[3] = {(double)(T)((double)a.y - di * kxc), (double)(T)((double)b.y - di * kyc),
                    (double)(T)((double)c.y - di * kzz)};
    if (idx == 0) {
        ukdc[0] = ur[0];
        ukdc[1] = ur[1];
        ukdc[2] = ur[2];
    }
    contra2co(gd, ur);
    contra2co(gd, ui);
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
    } else if (kz != 0) {
        c0 = 1;
        c1 = 0;
    } else {
        c0 = 0;
        c1 = 2;
    }
    z0[idx] = cmk((T)ore[c0], (T)oim[c0]);
    z1[idx] = cmk((T)ore[c1], (T)oim[c1]);
}

// ------------------------------------------------------------ launchers
template <typename T>
hipError_t obl_zeta_to_u(const Grid &g, cpx<T> *z0, cpx<T> *z1, cpx<T> *u0, cpx<T> *u1, cpx<T> *u2, cpx<T> *w1,
                         double *ukdc, int fixmask, int truncate, double gd, hipStream_t s) {
    const long long n = truncate == 0 ? sh_nkept(g) : g.nspec;
    hipLaunchKernelGGL(k_obl_zeta_to_u<T>, dim3(sh_blocks(n)), dim3(SH_NT), 0, s, g, z0, z1, u0, u1, u2, w1, ukdc,
                       fixmask, truncate, gd, (unsigned)n);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_zeta_to_omega(const Grid &g, const cpx<T> *z0, const cpx<T> *z1, cpx<T> *const o[3], hipStream_t s) {
    const long long n = sh_nkept(g);
    hipLaunchKernelGGL(k_obl_zeta_to_omega<T>, dim3(sh_blocks(n)), dim3(SH_NT), 0, s, g, z0, z1, o[0], o[1], o[2],
                       (unsigned)n);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_cross(long long n, T *const u[3], T *const w[3], hipStream_t s) {
    hipLaunchKernelGGL(k_obl_cross<T>, dim3(sh_blocks(n)), dim3(SH_NT), 0, s, n, u[0], u[1], u[2], w[0], w[1], w[2]);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_adv_ns(const Grid &g, cpx<T> *const F[3], const cpx<T> *w1, cpx<T> *z0, cpx<T> *z1, double rate,
                      double gd, double dmy0, double dmy1, hipStream_t s) {
    const long long nk = sh_nkept(g);
    hipLaunchKernelGGL(k_obl_adv_ns<T>, dim3(sh_blocks(nk)), dim3(SH_NT), 0, s, g, F[0], F[1], F[2], w1, z0, z1, rate,
                       gd, dmy0, dmy1, (unsigned)nk);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_remap(const Grid &g, T *const u[3], T *const o[3], double delta, hipStream_t s) {
    hipLaunchKernelGGL(k_obl_remap<T>, dim3(sh_blocks(g.nreal)), dim3(SH_NT), 0, s, g, u[0], u[1], u[2], o[0], o[1],
                       o[2], delta);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_to_cart(const Grid &g, T *const u[3], T *const c[3], double gd, double rate, const T *mask,
                       double *acc, double dx, hipStream_t s) {
    hipLaunchKernelGGL(k_obl_to_cart<T>, dim3(sh_row_blocks(g)), dim3(SH_NT), 0, s, g, u[0], u[1], u[2], c[0], c[1],
                       c[2], gd, rate, mask, acc, dx);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_add_f(const Grid &g, T *const up[3], const T *phisum, T *const c[3], T *const u[3], double gd,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_obl_add_f<T>, dim3(sh_row_blocks(g)), dim3(SH_NT), 0, s, g, up[0], up[1], up[2], phisum,
                       c[0], c[1], c[2], u[0], u[1], u[2], gd);
    return hipGetLastError();
}
template <typename T>
hipError_t obl_u_to_zeta(const Grid &g, const cpx<T> *u0, const cpx<T> *u1, const cpx<T> *u2, cpx<T> *z0, cpx<T> *z1,
                         double *ukdc, double gd, hipStream_t s, int kept) {
    const long long n = kept ? sh_nkept(g) : g.nspec;
    hipLaunchKernelGGL(k_obl_u_to_zeta<T>, dim3(sh_blocks(n)), dim3(SH_NT), 0, s, g, u0, u1, u2, z0, z1, ukdc, gd,
                       kept, (unsigned)n);
    return hipGetLastError();
}

#define KH_INST(T)                                                                                                   \
    template hipError_t obl_zeta_to_u<T>(const Grid &, cpx<T> *, cpx<T> *, cpx<T> *, cpx<T> *, cpx<T> *, cpx<T> *,   \
                                         double *, int, int, double, hipStream_t);                                   \
    template hipError_t obl_zeta_to_omega<T>(const Grid &, const cpx<T> *, const cpx<T> *, cpx<T> *const[3],         \
                                             hipStream_t);                                                           \
    template hipError_t obl_cross<T>(long long, T *const[3], T *const[3], hipStream_t);                              \
    template hipError_t obl_adv_ns<T>(const Grid &, cpx<T> *const[3], const cpx<T> *, cpx<T> *, cpx<T> *, double,    \
                                      double, double, double, hipStream_t);                                          \
    template hipError_t obl_remap<T>(const Grid &, T *const[3], T *const[3], double, hipStream_t);                   \
    template hipError_t obl_to_cart<T>(const Grid &, T *const[3], T *const[3], double, double, const T *,         \
                                       double *, double, hipStream_t);                                               \
    template hipError_t obl_add_f<T>(const Grid &, T *const[3], const T *, T *const[3], T *const[3], double,         \
                                     hipStream_t);                                                                   \
    template hipError_t obl_u_to_zeta<T>(const Grid &, const cpx<T> *, const cpx<T> *, const cpx<T> *, cpx<T> *,     \
                                         cpx<T> *, double *, double, hipStream_t, int);
KH_INST(double)
KH_INST(float)
#undef KH_INST

}  // namespace kh
