// kh_particle.hip -- particle <-> grid coupling and particle dynamics.
//
// Scatter and gather kernels run one workgroup per particle over that
// particle's Sekibun cell (the grid points within a + xi/2 + sqrt(3) dx of
// its grid cell, profile.cxx:22-65): each lane takes every 256th point,
// evaluates the profile in FP64 and either adds into the shared field with a
// hardware atomic (points where phi == 0 are skipped: adding +0 changes
// nothing) or accumulates its part of the force/torque sum. Sums are FP64
// with Neumaier compensation per lane, reduced across the workgroup with
// 64-wide wavefront shuffles and LDS. Integrators run one lane per particle,
// in FP64, for either field precision.
#include "kh_launch.h"

namespace kh {

constexpr int PT_NT = 256;  // lanes per particle workgroup (4 wavefronts)

struct Neu {  // Neumaier-compensated sum
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double val() const { return s + c; }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Reduce K doubles over the workgroup; result valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum(double *v) {
    __shared__ double red[PT_NT / 64][K];
    const int         lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double w = wave_sum(v[k]);
        if (lane == 0) red[wave][k] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            double t = 0.0;
            for (int w = 0; w < PT_NT / 64; w++) t += red[w][k];
            v[k] = t;
        }
    }
}

__device__ __forceinline__ void angular2v(int rotation, const double *om, const double *r, double *v) {
    if (rotation) {  // Angular2v_rot_on (make_phi.h:401-405)
        v[0] = om[1] * r[2] - om[2] * r[1];
        v[1] = om[2] * r[0] - om[0] * r[2];
        v[2] = om[0] * r[1] - om[1] * r[0];
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

// Make_phi_particle_sum scatter (make_phi.cxx:220-246). Also stores each
// point's phi and grid index in the per-particle cache the gather and the
// u_p scatter of the same step read (positions do not change in between).
// SCATTER = false only fills the cache.
template <typename T, bool SCATTER>
__global__ __launch_bounds__(PT_NT) void k_phi_scatter(PParams pp, const DevParticle *__restrict__ P,
                                                       T *__restrict__ phisum, T *__restrict__ cphi,
                                                       int *__restrict__ cidx) {
    const int       n     = blockIdx.x;
    const double    xp[3] = {P[n].x[0], P[n].x[1], P[n].x[2]};
    const PCell     c     = particle_cell(pp, xp);
    const long long cb    = (long long)n * pp.np_domain;
    for (int m = threadIdx.x; m < pp.np_domain; m += PT_NT) {
        const PPoint q  = sek_point(pp, c, xp, pp.sek[m]);
        const double ph = prof_phi(pp, q.dist);
        cphi[cb + m]    = (T)ph;
        cidx[cb + m]    = (int)q.idx;
        if (SCATTER && ph != 0.0) unsafeAtomicAdd(&phisum[q.idx], (T)ph);
    }
}

// r = cell*dx - residue (Relative_coord, make_phi.h:154-170)
__device__ __forceinline__ void sek_r(const PParams &pp, const PCell &c, int code, double *r) {
    int cell[3];
    sek_unpack(code, cell);
#pragma unroll
    for (int d = 0; d < 3; d++) r[d] = (double)cell[d] * pp.dx - c.residue[d];
}

// Calc_f_hydro_correct_precision (md_force.cxx:383-506), non-rigid branch,
// over the cached phi and grid index of each Sekibun point.
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_f_hydro(PParams pp, DevParticle *__restrict__ P, const T *__restrict__ u0,
                                                   const T *__restrict__ u1, const T *__restrict__ u2,
                                                   const T *__restrict__ phisum, const T *__restrict__ cphi,
                                                   const int *__restrict__ cidx, double coef) {
    const int       n     = blockIdx.x;
    const double    xp[3] = {P[n].x[0], P[n].x[1], P[n].x[2]};
    const double    vp[3] = {P[n].v[0], P[n].v[1], P[n].v[2]};
    const double    om[3] = {P[n].om[0], P[n].om[1], P[n].om[2]};
    const PCell     c     = particle_cell(pp, xp);
    const long long cb    = (long long)n * pp.np_domain;
    Neu             acc[6];
    for (int m = threadIdx.x; m < pp.np_domain; m += PT_NT) {
        const double ph = (double)cphi[cb + m];
        if (ph == 0.0) continue;
        const int idx = cidx[cb + m];
        double    r[3], vr[3];
        sek_r(pp, c, pp.sek[m], r);
        angular2v(pp.rotation, om, r, vr);
        const double dmy_phi = ph / fmax((double)phisum[idx], 1.0);
        double       fp[3];
        fp[0] = ((vp[0] + vr[0]) - (double)u0[idx]) * dmy_phi;
        fp[1] = ((vp[1] + vr[1]) - (double)u1[idx]) * dmy_phi;
        fp[2] = ((vp[2] + vr[2]) - (double)u2[idx]) * dmy_phi;
        acc[0].add(fp[0]);
        acc[1].add(fp[1]);
        acc[2].add(fp[2]);
        acc[3].add(r[1] * fp[2] - r[2] * fp[1]);
        acc[4].add(r[2] * fp[0] - r[0] * fp[2]);
        acc[5].add(r[0] * fp[1] - r[1] * fp[0]);
    }
    double v[6];
    for (int k = 0; k < 6; k++) v[k] = acc[k].val();
    block_sum<6>(v);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) {
            P[n].fh[d] = coef * v[d];
            P[n].fs[d] = 0.0;
        }
        if (pp.rotation)
            for (int d = 0; d < 3; d++) {
                P[n].th[d] = coef * v[3 + d];
                P[n].ts[d] = 0.0;
            }
    }
}

// Make_u_particle_sum scatter (make_phi.cxx:313-354) over the cache
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_up_scatter(PParams pp, const DevParticle *__restrict__ P,
                                                      const T *__restrict__ phisum, const T *__restrict__ cphi,
                                                      const int *__restrict__ cidx, T *__restrict__ up0,
                                                      T *__restrict__ up1, T *__restrict__ up2) {
    const int       n     = blockIdx.x;
    const double    xp[3] = {P[n].x[0], P[n].x[1], P[n].x[2]};
    const double    vp[3] = {P[n].v[0], P[n].v[1], P[n].v[2]};
    const double    om[3] = {P[n].om[0], P[n].om[1], P[n].om[2]};
    const PCell     c     = particle_cell(pp, xp);
    const long long cb    = (long long)n * pp.np_domain;
    for (int m = threadIdx.x; m < pp.np_domain; m += PT_NT) {
        const double ph = (double)cphi[cb + m];
        if (ph == 0.0) continue;
        const int    idx     = cidx[cb + m];
        const double dmy_phi = ph / fmax((double)phisum[idx], 1.0);
        double       r[3], vr[3];
        sek_r(pp, c, pp.sek[m], r);
        angular2v(pp.rotation, om, r, vr);
        unsafeAtomicAdd(&up0[idx], (T)((vp[0] + vr[0]) * dmy_phi));
        unsafeAtomicAdd(&up1[idx], (T)((vp[1] + vr[1]) * dmy_phi));
        unsafeAtomicAdd(&up2[idx], (T)((vp[2] + vr[2]) * dmy_phi));
    }
}

// ---------------------------------------------------------- quaternions
// qtn_prod (quaternion.h:160-168) with alpha; qdot SPACE_FRAME (rigid_body.cxx:157-168)
__device__ __forceinline__ void qdot_space(double *dq, const double *q, const double *w) {
    // qa = (0, w), qb = q, alpha = 0.5
    dq[0] = 0.5 * (0.0 * q[0] - (w[0] * q[1] + w[1] * q[2] + w[2] * q[3]));
    dq[1] = 0.5 * (0.0 * q[1] + q[0] * w[0] + w[1] * q[3] - w[2] * q[2]);
    dq[2] = 0.5 * (0.0 * q[2] + q[0] * w[1] + w[2] * q[1] - w[0] * q[3]);
    dq[3] = 0.5 * (0.0 * q[3] + q[0] * w[2] + w[0] * q[2] - w[1] * q[1]);
}
__device__ __forceinline__ void qtn_normalize(double *q) {  // quaternion.h:260-268
    double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    qn        = 1.0 / qn;
    for (int k = 0; k < 4; k++) q[k] *= qn;
}
// qtn_qvq_prod (quaternion.h:178-185): BODY2SPACE rotation
__device__ __forceinline__ void body2space(double *o, const double *q, const double *v) {
    const double qs2   = q[0] * q[0] - (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const double qdotp = q[1] * v[0] + q[2] * v[1] + q[3] * v[2];
    o[0]               = (qs2 * v[0] + 2.0 * (qdotp * q[1] + q[0] * (q[2] * v[2] - q[3] * v[1])));
    o[1]               = (qs2 * v[1] + 2.0 * (qdotp * q[2] + q[0] * (q[3] * v[0] - q[1] * v[2])));
    o[2]               = (qs2 * v[2] + 2.0 * (qdotp * q[3] + q[0] * (q[1] * v[1] - q[2] * v[0])));
}

// MD_solver_position_Euler / _AB2 (particle_solver.cxx:39-85) with the
// orientation updates (particle_rotation_solver.h:23-29, 72-81).
__global__ void k_position(PParams pp, DevParticle *__restrict__ P, int ab2, double dt, double hdt) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pp.np) return;
    DevParticle &p = P[n];
    if (pp.species[p.spec].propulsion == 2 /* obstacle */) return;
    for (int d = 0; d < 3; d++) {
        p.x_prev[d]          = p.x[d];
        const double delta_x = ab2 ? hdt * (3.0 * p.v[d] - p.v_old[d]) : dt * p.v[d];
        p.x_nopbc[d] += delta_x;
        p.x[d] += delta_x;
    }
    for (int d = 0; d < 3; d++) p.x[d] = fmod(p.x[d] + pp.L[d], pp.L[d]);
    if (pp.rotation) {
        double dq[4];
        if (!ab2) {
            for (int k = 0; k < 4; k++) p.q_old[k] = p.q[k];
            qdot_space(dq, p.q, p.om);
            for (int k = 0; k < 4; k++) p.q[k] += dq[k] * dt;
        } else {
            double dqo[4];
            qdot_space(dq, p.q, p.om);
            qdot_space(dqo, p.q_old, p.om_old);
            for (int k = 0; k < 4; k++) p.q_old[k] = p.q[k];
            for (int k = 0; k < 4; k++) p.q[k] += dq[k] * (3.0 * hdt);
            for (int k = 0; k < 4; k++) p.q[k] += dqo[k] * (-hdt);
        }
        qtn_normalize(p.q);
    }
}

// Lennard_Jones_f (interaction.h:160-199), powers 12:6, 24:12, 36:18
__device__ __forceinline__ double lj_f(const PParams &pp, double x) {
    double dmy = pp.lj_dia / x, ans = 0.0;
    if (pp.lj_powers == 0) {
        dmy = (dmy * dmy) * (dmy * dmy) * (dmy * dmy);
        ans = (24. * pp.epsilon) / (x * x) * (2.0 * (dmy * dmy) - dmy);
    } else if (pp.lj_powers == 1) {
        const double s = dmy * dmy;
        dmy            = s * s * s * s * s * s;
        ans            = (48. * pp.epsilon) / (x * x) * (2.0 * (dmy * dmy) - dmy);
    } else if (pp.lj_powers == 2) {
        const double s = dmy * dmy;
        dmy            = s * s * s * s * s * s * s * s * s;
        ans            = (72. * pp.epsilon) / (x * x) * (2.0 * (dmy * dmy) - dmy);
    }
    return ans;
}

// How many times the 16-unit link cell (md_force.cxx:14, 56-101) visits a
// pair: per axis 3 with one cell, 2 with two cells and different cells, else
// 1 for neighbouring cells and 0 otherwise.
__device__ __forceinline__ int lj_mult(const PParams &pp, const double *xi, const double *xj) {
    if (!pp.lj_cell_compat) return 1;
    const int N[3] = {pp.nx, pp.ny, pp.nz};
    int       m    = 1;
    for (int d = 0; d < 3; d++) {
        const int    lc = N[d] / 16;
        const double cl = 16 * pp.dx;
        const int    ci = (int)(xi[d] / cl), cj = (int)(xj[d] / cl);
        int          md;
        if (lc <= 1)
            md = 3;
        else if (lc == 2)
            md = (ci == cj) ? 1 : 2;
        else {
            const int dd = ((ci - cj) % lc + lc) % lc;
            md           = (dd == 0 || dd == 1 || dd == lc - 1) ? 1 : 0;
        }
        m *= md;
    }
    return m;
}

// Force (particle_solver.h:88-111): LJ pairs (md_force.cxx:34-196), gravity
// (md_force.cxx:306-324). One workgroup per particle i: its lanes split the
// partners j, each pair giving F_i += f(r_ij) (x_i - x_j), the same value the
// reference's +/- updates add; the lane sums are reduced in FP64.
__global__ __launch_bounds__(PT_NT) void k_forces(PParams pp, DevParticle *__restrict__ P) {
    const int    i      = blockIdx.x;
    const double xi[3]  = {P[i].x[0], P[i].x[1], P[i].x[2]};
    const int    si     = P[i].spec;
    const bool   obst_i = pp.species[si].propulsion == 2;
    double       f[3]   = {0.0, 0.0, 0.0};
    if (pp.lj_on) {
        for (int j = threadIdx.x; j < pp.np; j += PT_NT) {
            if (j == i) continue;
            if (obst_i && pp.species[P[j].spec].propulsion == 2) continue;
            const double xj[3] = {P[j].x[0], P[j].x[1], P[j].x[2]};
            double       rv[3];
            // r_ij_vec from the higher index to the lower, as the link cell takes it
            const double rij = (i > j) ? distance0(pp, xi, xj, rv) : distance0(pp, xj, xi, rv);
            if (rij < pp.lj_cutoff) {
                const int    mult  = lj_mult(pp, xi, xj);
                const double dmy_r = fmin(1.7976931348623157e308 / rij, lj_f(pp, rij));
                const double sgn   = (i > j) ? -1.0 : 1.0;  // i>j: f_i += dmy_r*(-r); else f_i -= dmy_r*(-r)
                for (int c = 0; c < mult; c++)
                    for (int d = 0; d < 3; d++) f[d] += sgn * (dmy_r * rv[d]);
            }
        }
    }
    block_sum<3>(f);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) P[i].fr[d] += f[d];
        if (pp.g != 0.0) P[i].fr[pp.g_dir] -= pp.gravity_on_fluid * (pp.species[si].mass_ratio - 1.0);
    }
}

// MD_solver_velocity_Euler (particle_solver.cxx:103-142) / _AB2_hydro
// (239-281), self_propulsion (88-101) and reset_Forces (11-37).
__global__ void k_velocity(PParams pp, DevParticle *__restrict__ P, int ab2, double h) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pp.np) return;
    DevParticle      &p = P[n];
    const DevSpecies &s = pp.species[p.spec];
    const double      dmy = h * s.imass, dmy_rot = h * s.imoi;
    if (s.propulsion != 2) {
        double sf[3] = {0, 0, 0}, st[3] = {0, 0, 0};
        if (pp.motor && s.propulsion == 0) {
            body2space(sf, p.q, s.janus_force);
            body2space(st, p.q, s.janus_torque);
        }
        for (int d = 0; d < 3; d++) {
            p.v_old[d]  = p.v[d];
            p.om_old[d] = p.om[d];
            if (!ab2) {
                p.v[d] += (dmy * (p.fh[d] + p.fr[d] + sf[d]));
                p.om[d] += (dmy_rot * (p.th[d] + p.tr[d] + st[d]));
            } else {
                p.v[d] += dmy * (2.0 * (p.fh[d] + sf[d]) + p.fr[d] + p.fr_prev[d]);
                p.om[d] += dmy_rot * (2.0 * (p.th[d] + st[d]) + p.tr[d] + p.tr_prev[d]);
            }
        }
    } else {
        for (int d = 0; d < 3; d++) {
            p.v_old[d]  = p.v[d];
            p.v[d]      = 0.0;
            p.om_old[d] = p.om[d];
            p.om[d]     = 0.0;
        }
    }
    for (int d = 0; d < 3; d++) {
        p.fr_prev[d] = p.fr[d];
        p.fr[d]      = 0.0;
        p.fh_prev[d] = p.fh[d];
        p.fh[d]      = 0.0;
        p.fs_prev[d] = p.fs[d];
        p.fs[d]      = 0.0;
        p.tr_prev[d] = p.tr[d];
        p.tr[d]      = 0.0;
        p.th_prev[d] = p.th[d];
        p.th[d]      = 0.0;
        p.ts_prev[d] = p.ts[d];
        p.ts[d]      = 0.0;
    }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}
__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo) {
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}

// Add_random_force_thermostat (fluct.cxx:10-43); normals from Philox keyed
// by (seed, stream) with counter (particle, step, block, stream).
__global__ void k_thermostat(PParams pp, DevParticle *__restrict__ P, uint64_t seed, uint32_t stream,
                             const int *__restrict__ d_ts, double niv, double nio) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pp.np) return;
    const int ts = *d_ts;
    DevParticle      &p = P[n];
    const DevSpecies &s = pp.species[p.spec];
    if (s.propulsion == 2) return;
    double g[6];
    for (int b = 0; b < 3; b++) {
        uint32_t c[4] = {(uint32_t)n, (uint32_t)ts, (uint32_t)b, stream};
        philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u0 = u53(c[0], c[1]), u1 = u53(c[2], c[3]);
        const double rad = sqrt(-2.0 * log(u0));
        const double ang = 2.0 * 3.14159265358979323846 * u1;
        g[2 * b]         = rad * cos(ang);
        g[2 * b + 1]     = rad * sin(ang);
    }
    for (int d = 0; d < 3; d++) {
        p.v[d] += g[d] * niv * s.imass;
        if (pp.rotation) p.om[d] += g[d + 3] * nio * s.imoi;
    }
}

// ------------------------------------------------------------ launchers
static inline unsigned pblocks(int n) { return (unsigned)((n + 63) / 64); }

template <typename T>
hipError_t phi_scatter(const PParams &pp, const DevParticle *P, T *phisum, T *cphi, int *cidx, int scatter,
                       hipStream_t s) {
    if (pp.np > 0) {
        if (scatter)
            hipLaunchKernelGGL((k_phi_scatter<T, true>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, phisum, cphi, cidx);
        else
            hipLaunchKernelGGL((k_phi_scatter<T, false>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, phisum, cphi, cidx);
    }
    return hipGetLastError();
}
template <typename T>
hipError_t f_hydro(const PParams &pp, DevParticle *P, T *const u[3], const T *phisum, const T *cphi, const int *cidx,
                   double coef, hipStream_t s) {
    if (pp.np > 0)
        hipLaunchKernelGGL(k_f_hydro<T>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P, u[0], u[1], u[2], phisum, cphi, cidx,
                           coef);
    return hipGetLastError();
}
template <typename T>
hipError_t up_scatter(const PParams &pp, const DevParticle *P, const T *phisum, const T *cphi, const int *cidx,
                      T *const up[3], hipStream_t s) {
    if (pp.np > 0)
        hipLaunchKernelGGL(k_up_scatter<T>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P, phisum, cphi, cidx, up[0], up[1],
                           up[2]);
    return hipGetLastError();
}
hipError_t position(const PParams &pp, DevParticle *P, int ab2, double dt, double hdt, hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_position, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, ab2, dt, hdt);
    return hipGetLastError();
}
hipError_t forces(const PParams &pp, DevParticle *P, hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_forces, dim3(pp.np), dim3(PT_NT), 0, s, pp, P);
    return hipGetLastError();
}
hipError_t velocity(const PParams &pp, DevParticle *P, int ab2, double h, hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_velocity, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, ab2, h);
    return hipGetLastError();
}
hipError_t thermostat(const PParams &pp, DevParticle *P, uint64_t seed, uint32_t stream, const int *d_ts, double niv,
                      double nio, hipStream_t s) {
    if (pp.np > 0)
        hipLaunchKernelGGL(k_thermostat, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, seed, stream, d_ts, niv, nio);
    return hipGetLastError();
}

// jikan.ts += 1 on the device (the step counter the noise stream is keyed by)
__global__ void k_tick(int *__restrict__ d_ts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *d_ts += 1;
}
hipError_t tick(int *d_ts, hipStream_t s) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, d_ts);
    return hipGetLastError();
}

#define KH_INST(T)                                                                                              \
    template hipError_t phi_scatter<T>(const PParams &, const DevParticle *, T *, T *, int *, int, hipStream_t);   \
    template hipError_t f_hydro<T>(const PParams &, DevParticle *, T *const[3], const T *, const T *, const int *, \
                                   double, hipStream_t);                                                          \
    template hipError_t up_scatter<T>(const PParams &, const DevParticle *, const T *, const T *, const int *,     \
                                      T *const[3], hipStream_t);
KH_INST(double)
KH_INST(float)
#undef KH_INST

}  // namespace kh
