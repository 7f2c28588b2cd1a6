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

// The fused particle step keeps, for the first PT_CAP cached points of its
// particle, what the force gather computed (phi/max(phi_sum,1), negated
// where phi_sum is this particle's phi alone, the grid index, the offset) in
// LDS for the u_p scatter, which then reads no global memory for them.
constexpr int PT_CAP = 1024;
template <typename T>
struct PStash {
    T   dmy[PT_CAP];
    int idx[PT_CAP];
    int code[PT_CAP];
};

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
template <int K, int NT = PT_NT>
__device__ __forceinline__ void block_sum(double *v) {
    __shared__ double red[NT / 64][K];
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
            for (int w = 0; w < NT / 64; w++) t += red[w][k];
            v[k] = t;
        }
    }
}

template <typename T>
__device__ __forceinline__ void angular2v_t(int rotation, const T *om, const T *r, T *v) {
    if (rotation) {
        v[0] = om[1] * r[2] - om[2] * r[1];
        v[1] = om[2] * r[0] - om[0] * r[2];
        v[2] = om[0] * r[1] - om[1] * r[0];
    } else {
        v[0] = v[1] = v[2] = (T)0;
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
// SCATTER = false only fills the cache. MOVE = true first advances the
// particle (the position integrator of the same step, one lane).
__device__ void position_one(const PParams &pp, DevParticle *P, int n, int ab2, double dt, double hdt);

template <typename T, bool SCATTER, bool MOVE, int NT, bool OBL>
__device__ __forceinline__ void phi_scatter_block(const PParams &pp, DevParticle *__restrict__ P,
                                                  T *__restrict__ phisum, const PCache<T> &pc, int ab2, double dt,
                                                  double hdt, int n) {
    // the first Sekibun offset of this lane, loaded before the position
    // update so its latency hides behind it; each round loads the next
    int next = threadIdx.x < pp.np_domain ? pp.sek[threadIdx.x] : 0;
    double    xp[3];
    if constexpr (MOVE) {
        __shared__ double xs[3];
        if (threadIdx.x == 0) {
            position_one(pp, P, n, ab2, dt, hdt);
            for (int d = 0; d < 3; d++) xs[d] = P[n].x[d];
        }
        __syncthreads();
        for (int d = 0; d < 3; d++) xp[d] = xs[d];
    } else {
        for (int d = 0; d < 3; d++) xp[d] = P[n].x[d];
    }
    const PCell     c  = particle_cell(pp, xp);
    const long long cb = (long long)n * pp.np_domain;
    // the nonzero points are packed in Sekibun order: per round of NT
    // points, a wave's ballot gives each lane its rank among the wave's
    // nonzero points and the per-wave counts (double-buffered in LDS) the
    // wave's base
    __shared__ int wcnt[2][NT / 64];
    const int      lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int            off  = 0;
    for (int m0 = 0, par = 0; m0 < pp.np_domain; m0 += NT, par ^= 1) {
        const int m    = m0 + threadIdx.x;
        T         ph   = (T)0;
        long long idx  = 0;
        const int code = next;
        if (m + NT < pp.np_domain) next = pp.sek[m + NT];
        if (m < pp.np_domain) {
            PPoint q;
            if (sizeof(T) == 4 && pp.direct_dist) {  // FP32 distance of the offset
                q              = sek_point<false, OBL>(pp, c, xp, code);
                const float r0 = (float)q.r[0], r1 = (float)q.r[1], r2 = (float)q.r[2];
                ph             = prof_phi_t<T>(pp, (T)sqrtf(r0 * r0 + r1 * r1 + r2 * r2));
            } else {
                q  = sek_point<true, OBL>(pp, c, xp, code);
                ph = prof_phi_t<T>(pp, (T)q.dist);
            }
            idx = q.idx;
            if (SCATTER && ph != (T)0) unsafeAtomicAdd(&phisum[idx], ph);
        }
        const bool               keep = ph != (T)0;
        if constexpr (SCATTER && !OBL && occ_kept<T>()) {
            if (pp.occ.w) {
                // mark the 4-point vectors this round reaches. Sekibun order
                // keeps a row's points on consecutive lanes and its nonzero
                // points form one chord, so the first lane of each run of
                // nonzero points in one row marks the run's whole range of
                // vectors (wrapping at the row's end) at once.
                const int  row  = (int)(idx / pp.nz);
                const int  v    = (int)(idx - (long long)row * pp.nz) >> 2;
                const int  prow = __shfl_up(row, 1), nrow = __shfl_down(row, 1);
                const int  pk = __shfl_up((int)keep, 1), nk = __shfl_down((int)keep, 1);
                const bool head = keep && (lane == 0 || !pk || prow != row);
                const bool tail = keep && (lane == 63 || !nk || nrow != row);
                const unsigned long long tails = __ballot(tail);
                const unsigned long long after = tails & (~0ull << lane);
                const int e  = after ? __ffsll((long long)after) - 1 : lane;  // the run's last lane
                const int ve = __shfl(v, e);
                if (head) {
                    const int nv = (pp.nz + 3) / 4;
                    if (v <= ve) {
                        occ_mark(pp.occ, row, v, ve);
                    } else {
                        occ_mark(pp.occ, row, v, nv - 1);
                        occ_mark(pp.occ, row, 0, ve);
                    }
                }
            }
        }
        const unsigned long long b    = __ballot(keep);
        if (lane == 0) wcnt[par][wave] = __popcll(b);
        __syncthreads();
        int base = off, tot = 0;
#pragma unroll
        for (int w = 0; w < NT / 64; w++) {
            const int k = wcnt[par][w];
            base += w < wave ? k : 0;
            tot += k;
        }
        if (keep) {
            const long long j = cb + base + __popcll(b & __lanemask_lt());
            pc.phi[j]         = ph;
            pc.idx[j]         = (int)idx;
            pc.code[j]        = code;
        }
        off += tot;
    }
    if (threadIdx.x == 0) pc.cnt[n] = off;
}
template <typename T, bool SCATTER, bool MOVE, int NT, bool OBL = false>
__global__ __launch_bounds__(NT) void k_phi_scatter(PParams pp, DevParticle *__restrict__ P,
                                                       T *__restrict__ phisum, PCache<T> pc, int ab2, double dt,
                                                       double hdt) {
    phi_scatter_block<T, SCATTER, MOVE, NT, OBL>(pp, P, phisum, pc, ab2, dt, hdt, blockIdx.x);
}

// r = cell*dx - residue (Relative_coord, make_phi.h:154-170); on a sheared
// run the stepover form, and the returned sign of the point's y crossing
template <bool OBL = false>
__device__ __forceinline__ int sek_r(const PParams &pp, const PCell &c, int code, double *r) {
    int cell[3];
    sek_unpack(code, cell);
#pragma unroll
    for (int d = 0; d < 3; d++) r[d] = (double)cell[d] * pp.dx - c.residue[d];
    if constexpr (!OBL) {
        return 0;
    } else {
        int rm0;
        return obl_stepover(pp, c, cell, &rm0, &r[0]);
    }
}

// Calc_f_hydro_correct_precision (md_force.cxx:383-506), non-rigid branch,
// over the cached phi and grid index of each Sekibun point.
template <typename T, int NT = PT_NT, bool OBL = false>
__device__ __forceinline__ void f_hydro_block(const PParams &pp, DevParticle *__restrict__ P, int n,
                                              const T *__restrict__ u0, const T *__restrict__ u1,
                                              const T *__restrict__ u2, const T *__restrict__ phisum,
                                              const PCache<T> &pc, double coef, PStash<T> *stash = nullptr) {
    const double    xp[3] = {P[n].x[0], P[n].x[1], P[n].x[2]};
    const double    vp[3] = {P[n].v[0], P[n].v[1], P[n].v[2]};
    const double    om[3] = {P[n].om[0], P[n].om[1], P[n].om[2]};
    const PCell     c     = particle_cell(pp, xp);
    const long long cb    = (long long)n * pp.np_domain;
    Neu             acc[6];
    // per-point terms in the field precision T, sums in FP64
    const T vt[3] = {(T)vp[0], (T)vp[1], (T)vp[2]}, wt[3] = {(T)om[0], (T)om[1], (T)om[2]};
    const int cnt = pc.cnt[n];
    // sheared runs: the Hydro_force sums (md_force.cxx:620-745) need, per
    // particle, sum fp_x y', sum w, sum w |r|^2, sum y' w, sum r_y y' w and
    // sum r_z y' w (w = phi/max(phi_sum,1), y' = r_mesh_y + sign LY)
    Neu st[6];
    for (int m = threadIdx.x; m < cnt; m += NT) {
        const T   ph   = pc.phi[cb + m];
        const int idx  = pc.idx[cb + m];
        const int code = pc.code[cb + m];
        double    rd[3];
        const int sgn = sek_r<OBL>(pp, c, code, rd);
        const T r[3] = {(T)rd[0], (T)rd[1], (T)rd[2]};
        T       vr[3];
        angular2v_t<T>(pp.rotation, wt, r, vr);
        // across the sheared boundary the image moves sgn*obl_jump faster
        // (md_force.cxx:614); the reference forms vp - sign*jump first
        T vx = vt[0];
        if constexpr (OBL) vx = sgn ? (T)(vp[0] - sgn * pp.obl_jump) : vt[0];
        const T ps      = phisum[idx];
        const T dmy_phi = ph / (ps > (T)1 ? ps : (T)1);
        if (stash && m < PT_CAP) {
            stash->dmy[m]  = ps == ph ? -dmy_phi : dmy_phi;
            stash->idx[m]  = idx;
            stash->code[m] = code;
        }
        T       fp[3];
        fp[0] = ((vx + vr[0]) - u0[idx]) * dmy_phi;
        fp[1] = ((vt[1] + vr[1]) - u1[idx]) * dmy_phi;
        fp[2] = ((vt[2] + vr[2]) - u2[idx]) * dmy_phi;
        acc[0].add((double)fp[0]);
        acc[1].add((double)fp[1]);
        acc[2].add((double)fp[2]);
        acc[3].add((double)(r[1] * fp[2] - r[2] * fp[1]));
        acc[4].add((double)(r[2] * fp[0] - r[0] * fp[2]));
        acc[5].add((double)(r[0] * fp[1] - r[1] * fp[0]));
        if constexpr (OBL) {
            const double ry = (double)((idx / pp.nz) % pp.ny) + sgn * pp.L[1], w = (double)dmy_phi;
            st[0].add((double)fp[0] * ry);
            st[1].add(w);
            st[2].add(w * (rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]));
            st[3].add(ry * w);
            st[4].add(rd[1] * ry * w);
            st[5].add(rd[2] * ry * w);
        }
    }
    constexpr int K = OBL ? 12 : 6;
    double        v[K];
    for (int k = 0; k < 6; k++) v[k] = acc[k].val();
    if constexpr (OBL)
        for (int k = 0; k < 6; k++) v[6 + k] = st[k].val();
    block_sum<K, NT>(v);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) {
            P[n].fh[d] = coef * v[d];
            if constexpr (!OBL) P[n].fs[d] = 0.0;
        }
        if (pp.rotation)
            for (int d = 0; d < 3; d++) {
                P[n].th[d] = coef * v[3 + d];
                if constexpr (!OBL) P[n].ts[d] = 0.0;
            }
        if constexpr (OBL) {
            // per-particle partials of Hydro_force (old) and Hydro_force_new;
            // the mean-force term and the sums over particles follow in
            // k_obl_record
            const double rhop = pp.species[P[n].spec].mass_ratio * pp.rho;
            const double a = rhop * v[6], vol = v[7], itr = 2.0 / 3.0 * v[8], b = v[9];
            double      *o = pp.pstress + 6 * (long long)n;
            o[0]           = a - (P[n].mdfr[0] / vol) * b;
            o[1]           = a + (-pp.rho * v[0] / vol) * b + (-pp.rho / itr) * (v[4] * v[11] - v[5] * v[10]);
            o[2]           = b;
            o[3]           = v[0];
            o[4]           = vol;
            o[5]           = 0.0;  // the LJ pairs' shear stress (k_forces)
        }
    }
}
template <typename T, bool OBL>
__global__ __launch_bounds__(PT_NT) void k_f_hydro(PParams pp, DevParticle *__restrict__ P, const T *__restrict__ u0,
                                                   const T *__restrict__ u1, const T *__restrict__ u2,
                                                   const T *__restrict__ phisum, PCache<T> pc, double coef) {
    f_hydro_block<T, PT_NT, OBL>(pp, P, blockIdx.x, u0, u1, u2, phisum, pc, coef);
}

// Make_u_particle_sum scatter (make_phi.cxx:313-354) over the cache. With
// u* given (us0..2) each particle also subtracts its share w u* of the
// rigidity term, w = phi/max(phi_sum,1): the shares sum to min(phi_sum,1)
// u*, so the field receives f = u_p - min(phi_sum,1) u*
// (Make_f_particle_dt_nonsole, f_particle.h:35-54) directly; where one
// particle alone covers a point (w = phi_sum = phi) the value is the
// reference's to the last bit.
template <typename T, int NT = PT_NT, bool OBL = false>
__device__ __forceinline__ void up_scatter_block(const PParams &pp, int n, const double *xp, const double *vp,
                                                 const double *om, const T *__restrict__ phisum,
                                                 const PCache<T> &pc, T *__restrict__ up0, T *__restrict__ up1,
                                                 T *__restrict__ up2, const PStash<T> *stash = nullptr,
                                                 const T *__restrict__ us0 = nullptr,
                                                 const T *__restrict__ us1 = nullptr,
                                                 const T *__restrict__ us2 = nullptr) {
    const PCell     c     = particle_cell(pp, xp);
    const long long cb    = (long long)n * pp.np_domain;
    const T vt[3] = {(T)vp[0], (T)vp[1], (T)vp[2]}, wt[3] = {(T)om[0], (T)om[1], (T)om[2]};
    const int cnt = pc.cnt[n];
    for (int m = threadIdx.x; m < cnt; m += NT) {
        int  idx, code;
        T    dmy_phi;
        bool alone;  // phi_sum is this particle's phi alone
        if (stash && m < PT_CAP) {
            const T d = stash->dmy[m];
            idx       = stash->idx[m];
            code      = stash->code[m];
            alone     = d < (T)0;
            dmy_phi   = alone ? -d : d;
        } else {
            const T ph = pc.phi[cb + m];
            idx        = pc.idx[cb + m];
            code       = pc.code[cb + m];
            const T ps = phisum[idx];
            dmy_phi    = ph / (ps > (T)1 ? ps : (T)1);
            alone      = ps == ph;
        }
        double    rd[3];
        const int sgn = sek_r<OBL>(pp, c, code, rd);
        const T r[3] = {(T)rd[0], (T)rd[1], (T)rd[2]};
        T       vr[3];
        angular2v_t<T>(pp.rotation, wt, r, vr);
        T vx = vt[0];
        if constexpr (OBL) vx = sgn ? (T)(vp[0] - sgn * pp.obl_jump) : vt[0];  // make_phi.cxx:392
        T a = (vx + vr[0]) * dmy_phi, b = (vt[1] + vr[1]) * dmy_phi, e = (vt[2] + vr[2]) * dmy_phi;
        if (us0) {
            a = a - dmy_phi * us0[idx];
            b = b - dmy_phi * us1[idx];
            e = e - dmy_phi * us2[idx];
        }
        if (alone) {
            // phi_sum is this particle's phi alone: no other particle adds
            // here (a contribution below half an ulp of phi would be the only
            // exception, and it is below the rounding of the sum), so the
            // sum onto the zeroed field is a plain store
            up0[idx] = a;
            up1[idx] = b;
            up2[idx] = e;
        } else {
            unsafeAtomicAdd(&up0[idx], a);
            unsafeAtomicAdd(&up1[idx], b);
            unsafeAtomicAdd(&up2[idx], e);
        }
    }
}
template <typename T, bool OBL>
__global__ __launch_bounds__(PT_NT) void k_up_scatter(PParams pp, const DevParticle *__restrict__ P,
                                                      const T *__restrict__ phisum, PCache<T> pc,
                                                      T *__restrict__ up0, T *__restrict__ up1,
                                                      T *__restrict__ up2) {
    const int    n     = blockIdx.x;
    const double xp[3] = {P[n].x[0], P[n].x[1], P[n].x[2]};
    const double vp[3] = {P[n].v[0], P[n].v[1], P[n].v[2]};
    const double om[3] = {P[n].om[0], P[n].om[1], P[n].om[2]};
    up_scatter_block<T, PT_NT, OBL>(pp, n, xp, vp, om, phisum, pc, up0, up1, up2);
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
    if (n < pp.np) position_one(pp, P, n, ab2, dt, hdt);
}
__device__ void position_one(const PParams &pp, DevParticle *P, int n, int ab2, double dt, double hdt) {
    DevParticle &p = P[n];
    if (pp.species[p.spec].propulsion == 2 /* obstacle */) return;
    for (int d = 0; d < 3; d++) {
        p.x_prev[d]          = p.x[d];
        const double delta_x = ab2 ? hdt * (3.0 * p.v[d] - p.v_old[d]) : dt * p.v[d];
        p.x_nopbc[d] += delta_x;
        p.x[d] += delta_x;
    }
    if (!pp.obl) {
        for (int d = 0; d < 3; d++) p.x[d] = fmod(p.x[d] + pp.L[d], pp.L[d]);
    } else {
        // PBC_OBL (periodic_boundary.h:26-44): through the top (bottom) the
        // particle enters the row of boxes below (above): x shifted by
        // -(+)strain*LY, v_x by -(+)rate*LY
        double signY = p.x[1];
        p.x[1]       = fmod(p.x[1] + pp.L[1], pp.L[1]);
        signY -= p.x[1];
        int sign = (int)signY;
        if (!(sign == 0)) sign = sign / abs(sign);
        p.x[0] -= (double)sign * pp.obl_g * pp.L[1];
        p.x[0]              = fmod(p.x[0] + pp.L[0], pp.L[0]);
        p.x[2]              = fmod(p.x[2] + pp.L[2], pp.L[2]);
        const double dvx    = -(double)sign * pp.obl_jump;
        p.v[0] += dvx;
        p.v_old[0] += dvx;
    }
    for (int d = 0; d < 3; d++) pp.xs[4 * n + d] = p.x[d];
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
// RIGID: beads of rigid bodies -- pairs within one body are skipped
// (rigid_chain, interaction.h:18-24), gravity acts on the bodies, and the
// forces from partners of higher index are also kept apart (pp.flo) for
// the reference's body torque bookkeeping.
template <int NT = PT_NT, bool OBL = false, bool RIGID = false>
__device__ __forceinline__ void forces_block(const PParams &pp, DevParticle *__restrict__ P, int i) {
    const double xi[3]  = {P[i].x[0], P[i].x[1], P[i].x[2]};
    const int    si     = P[i].spec;
    const bool   obst_i = pp.species[si].propulsion == 2;
    // F_i; [3]: sheared runs, the pairs' shear stress; rigid beads: [3..5]
    // the part from partners of higher index
    double f[RIGID ? 6 : 4] = {};
    const int bi = RIGID ? pp.bid[i] : -1;
    if (pp.lj_on) {
        // pairs clearly beyond the cutoff are dropped by a division-free
        // minimum image (exact for positions inside the box, up to the
        // rounding of ties at half a box, hence the margin on the cutoff);
        // the rest take the reference's Distance0
        const double cut2 = pp.lj_cutoff * pp.lj_cutoff * (1.0 + 1e-9);
        for (int j = threadIdx.x; j < pp.np; j += NT) {
            if (j == i) continue;
            if (obst_i && pp.xs[4 * j + 3] != 0.0) continue;
            if (RIGID && pp.bid[j] == bi) continue;
            const double xj[3] = {pp.xs[4 * j], pp.xs[4 * j + 1], pp.xs[4 * j + 2]};
            double       d2    = 0.0;
            bool         odd   = false;  // a position outside the box
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double hl = 0.5 * pp.L[d];
                double       e  = xj[d] - xi[d];
                e               = e > hl ? e - pp.L[d] : (e < -hl ? e + pp.L[d] : e);
                odd |= fabs(e) > hl;
                d2 += e * e;
            }
            if (!OBL && !odd && d2 >= cut2) continue;
            double rv[3];
            double rij;
            if constexpr (!OBL) {
                // r_ij_vec from the higher index to the lower, as the link cell takes it
                rij = (i > j) ? distance0(pp, xi, xj, rv) : distance0(pp, xj, xi, rv);
            } else {
                // Calc_f_Lennard_Jones_shear_cap_primitive (md_force.cxx:198-305)
                // takes r = Distance0_OBL(x_n, x_m) for n < m (x_m - x_n); its
                // negation has the orientation the update below expects
                rij = (i < j) ? distance0_obl(pp, xi, xj, rv) : distance0_obl(pp, xj, xi, rv);
                for (int d = 0; d < 3; d++) rv[d] = -rv[d];
            }
            if (rij < pp.lj_cutoff) {
                const int    mult  = lj_mult(pp, xi, xj);
                const double dmy_r = fmin(1.7976931348623157e308 / rij, lj_f(pp, rij));
                const double sgn   = (i > j) ? -1.0 : 1.0;  // i>j: f_i += dmy_r*(-r); else f_i -= dmy_r*(-r)
                for (int c = 0; c < mult; c++)
                    for (int d = 0; d < 3; d++) {
                        f[d] += sgn * (dmy_r * rv[d]);
                        if (RIGID && j > i) f[3 + d] += sgn * (dmy_r * rv[d]);
                    }
                // md_force.cxx:246: F_x on the lower index times r_y, each pair
                // once (r = x_hi - x_lo = -rv here)
                if constexpr (OBL)
                    if (i < j) f[3] += (dmy_r * rv[0]) * (-rv[1]);
            }
        }
    }
    block_sum<RIGID ? 6 : (OBL ? 4 : 3), NT>(f);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) P[i].fr[d] += f[d];
        if constexpr (RIGID) {
            for (int d = 0; d < 3; d++) pp.flo[3 * (long long)i + d] = f[3 + d];
        } else if (pp.g != 0.0) {
            P[i].fr[pp.g_dir] -= pp.gravity_on_fluid * (pp.species[si].mass_ratio - 1.0);
        }
        if constexpr (OBL) pp.pstress[6 * (long long)i + 5] = f[3];
    }
}
template <bool OBL, bool RIGID = false>
__global__ __launch_bounds__(PT_NT) void k_forces(PParams pp, DevParticle *__restrict__ P) {
    forces_block<PT_NT, OBL, RIGID>(pp, P, blockIdx.x);
}

// The record Mean_shear_stress prints after a sheared step (sp_3d_ns.h:
// 247-330): the particle sums of Hydro_force / Hydro_force_new with the
// mean-force term (md_force.cxx:681-745) as Calc_hydro_stress takes them
// (fluid_solver.h:253-279), the LJ stress over the volume (Force_OBL,
// particle_solver.h:113-126), the Reynolds stress from k_obl_to_cart's
// per-workgroup partials. One workgroup.
__global__ __launch_bounds__(PT_NT) void k_obl_record(PParams pp, const double *__restrict__ acc, long long nacc,
                                                      ShearRec h, double *__restrict__ last,
                                                      double *__restrict__ hist) {
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int n = threadIdx.x; n < pp.np; n += PT_NT)
        for (int k = 0; k < 6; k++) v[k] += pp.pstress[6 * (long long)n + k];
    // eight running sums keep eight loads of the partials in flight
    double    a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long b    = threadIdx.x;
    for (; b + 7 * PT_NT < nacc; b += 8 * PT_NT)
        for (int u = 0; u < 8; u++) a[u] += acc[b + u * PT_NT];
    for (; b < nacc; b += PT_NT) a[0] += acc[b];
    v[6] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    block_sum<7, PT_NT>(v);
    if (threadIdx.x == 0) {
        // (no particles: the mean-force spread has no points, md_force.cxx:719-745)
        const double sum_volume = v[4] / pp.rho;
        const double old        = pp.np > 0 ? v[0] - (v[3] / sum_volume) * v[2] : v[0];
        double       rec[12];
        rec[0]  = h.time;
        rec[1]  = h.rate;
        rec[2]  = h.deg;
        rec[3]  = h.strain;
        rec[4]  = v[5] * h.ivolume;                       // lj_dev_stress
        rec[5]  = -old * h.ivol_dx3 / h.dt;               // shear_stress_temporal_old
        rec[6]  = -v[1] * h.ivol_dx3 / h.dt;              // shear_stress_temporal_new
        rec[7]  = -h.ivol_dx3 * pp.rho * v[6];            // reynolds_stress
        rec[8]  = h.eta * h.rate;                         // fluid_stress
        rec[9]  = 0.0;                                    // interfacial_stress
        rec[10] = rec[6] + rec[7] + rec[4] + rec[8] + rec[9];  // apparent_stress
        rec[11] = rec[10] / h.rate;                       // viscosity
        for (int k = 0; k < 12; k++) last[k] = rec[k];
        if (hist)
            for (int k = 0; k < 12; k++) hist[k] = rec[k];
    }
}
hipError_t obl_record(const PParams &pp, const double *acc, long long nacc, const ShearRec &h, double *last,
                      double *hist, hipStream_t s) {
    hipLaunchKernelGGL(k_obl_record, dim3(1), dim3(PT_NT), 0, s, pp, acc, nacc, h, last, hist);
    return hipGetLastError();
}

// MD_solver_velocity_Euler (particle_solver.cxx:103-142) / _AB2_hydro
// (239-281), self_propulsion (88-101) and reset_Forces (11-37).
__device__ void add_noise(const PParams &pp, DevParticle *P, int n, const Noise &nz);

__device__ void velocity_one(const PParams &pp, DevParticle *__restrict__ P, int n, int ab2, double h,
                             const Noise &nz);
__global__ void k_velocity(PParams pp, DevParticle *__restrict__ P, int ab2, double h, Noise nz) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < pp.np) velocity_one(pp, P, n, ab2, h, nz);
}
__device__ void velocity_one(const PParams &pp, DevParticle *__restrict__ P, int n, int ab2, double h,
                             const Noise &nz) {
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
    // hydro_stress of the sheared velocity updates (particle_solver.cxx:336-432):
    // the LJ + gravity impulse of this update
    if (pp.obl)
        for (int d = 0; d < 3; d++) p.mdfr[d] = ab2 ? h * (p.fr[d] + p.fr_prev[d]) : h * p.fr[d];
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
    if (nz.on) add_noise(pp, P, n, nz);  // the thermostat follows the update (sp_3d_ns.cxx)
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
__global__ void k_thermostat(PParams pp, DevParticle *__restrict__ P, Noise nz) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < pp.np) add_noise(pp, P, n, nz);
}
__device__ void add_noise(const PParams &pp, DevParticle *P, int n, const Noise &nz) {
    const uint64_t seed = nz.seed;
    const uint32_t stream = nz.stream;
    const double   niv = nz.niv, nio = nz.nio;
    const int      ts = *nz.d_ts;
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

// The particle part of the fused step after the phi scatter, one workgroup
// per particle: Calc_f_hydro_correct_precision, Force (LJ + gravity), the
// velocity update with the thermostat's kick, Make_u_particle_sum. Each
// stage reads only this particle's state and inputs already complete when
// the launch starts (u*, phi_sum, every particle's new position), so the
// four reference loops need no grid-wide barrier between them.
// NT = 256: 8 waves per SIMD (64 VGPRs), so two thousand particle
// workgroups are all resident at once on 256 CUs. Few particles leave the
// chip idle and the step's time is one workgroup's chain of dependent
// loads: NT = 1024 splits each particle's points over four times the lanes.
template <typename T, int NT>
__device__ __forceinline__ void particle_step_block(const PParams &pp, DevParticle *__restrict__ P,
                                                    const T *__restrict__ u0, const T *__restrict__ u1,
                                                    const T *__restrict__ u2, const T *__restrict__ phisum,
                                                    const PCache<T> &pc, double coef, int ab2, double h,
                                                    const Noise &nz, T *__restrict__ up0, T *__restrict__ up1,
                                                    T *__restrict__ up2, int n) {
    __shared__ PStash<T>    stash;
    f_hydro_block<T, NT>(pp, P, n, u0, u1, u2, phisum, pc, coef, &stash);
    forces_block<NT>(pp, P, n);
    __shared__ double st[9];
    if (threadIdx.x == 0) {
        velocity_one(pp, P, n, ab2, h, nz);
        for (int d = 0; d < 3; d++) {
            st[d]     = P[n].x[d];
            st[3 + d] = P[n].v[d];
            st[6 + d] = P[n].om[d];
        }
    }
    __syncthreads();
    // each lane reads back only the stash entries it wrote; the field
    // receives f = u_p - min(phi_sum,1) u* (z_force then reads it as it is)
    up_scatter_block<T, NT>(pp, n, st, st + 3, st + 6, phisum, pc, up0, up1, up2, &stash, u0, u1, u2);
}
template <typename T, int NT>
__global__ __launch_bounds__(NT, NT == PT_NT ? 8 : 1) void k_particle_step(PParams pp, DevParticle *__restrict__ P,
                                                         const T *__restrict__ u0, const T *__restrict__ u1,
                                                         const T *__restrict__ u2, const T *__restrict__ phisum,
                                                         PCache<T> pc, double coef, int ab2, double h, Noise nz,
                                                         T *__restrict__ up0, T *__restrict__ up1,
                                                         T *__restrict__ up2) {
    particle_step_block<T, NT>(pp, P, u0, u1, u2, phisum, pc, coef, ab2, h, nz, up0, up1, up2, blockIdx.x);
}

// The same two particle launches of a steady-state step for a batch of runs
// (kh_batch): grid z is the slot, each slot's arguments come from its
// SlotPart record, slots with active[z] == 0 and blocks past the slot's
// particle count return.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_phi_scatter_batch(const SlotPart<T> *__restrict__ sp,
                                                          const int *__restrict__ active, double dt, double hdt) {
    if (!active[blockIdx.z]) return;
    const SlotPart<T> &a = sp[blockIdx.z];
    if ((int)blockIdx.x >= a.pp.np) return;
    phi_scatter_block<T, true, true, NT, false>(a.pp, a.P, a.phisum, a.pc, 1, dt, hdt, blockIdx.x);
}
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_particle_step_batch(const SlotPart<T> *__restrict__ sp,
                                                            const int *__restrict__ active, double coef, double h) {
    if (!active[blockIdx.z]) return;
    const SlotPart<T> &a = sp[blockIdx.z];
    if ((int)blockIdx.x >= a.pp.np) return;
    particle_step_block<T, NT>(a.pp, a.P, a.u[0], a.u[1], a.u[2], a.phisum, a.pc, coef, 1, h, a.nz, a.up[0], a.up[1],
                               a.up[2], blockIdx.x);
}

// ------------------------------------------------------------ squirmers
// The slip iteration of Time_evolution_hydro (sp_3d_ns.cxx:121-161) for
// SW_JANUS_SLIP species: surface slip u_s = B1 (sin th + (B2/2B1) sin 2th) e_th
// + (C1 sin th + 1.5 C2 sin 2th) e_ph in the thin shell |r - a| < xi/2,
// balanced by a droplet counter-flow so the particle's momentum is unchanged.

__device__ __forceinline__ bool less_than_mp(double a, double b) {  // macro.h:164-167
    const double eps = (fmax(fabs(a), fabs(b)) + 10.0) * 2.220446049250313080847e-16;
    return (b - a) > eps;
}
// DPhi_compact_sin_norm (profile.h:175-184)
__device__ __forceinline__ double dphi_sin(const PParams &pp, double x) {
    const double dmy = pp.radius - x;
    return (fabs(dmy) < pp.hxi) ? cos(3.14159265358979323846 * dmy / pp.xi) : 0.0;
}
// SPACE2BODY rotation: qvq with the conjugate quaternion
__device__ __forceinline__ void space2body(double *o, const double *q, const double *v) {
    const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
    body2space(o, qc, v);
}
__device__ __forceinline__ void vcross(double *c, const double *a, const double *b) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Squirmer_coord (make_phi.h:282-376): unit vectors e_r, e_th, e_ph (space
// frame) and the polar angle th of the body-frame point x about the
// species' Janus axis.
__device__ __forceinline__ void squirmer_coord(const double *x, const double *q, int axis, double *nr, double *nth,
                                               double *nph, double *theta) {
    double e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0}, e3[3] = {0, 0, 0};
    if (axis == 2) {  // z
        e1[0] = 1;
        e2[1] = 1;
        e3[2] = 1;
    } else if (axis == 1) {  // y
        e1[2] = 1;
        e2[0] = 1;
        e3[1] = 1;
    } else {  // x
        e1[1] = 1;
        e2[2] = 1;
        e3[0] = 1;
    }
    double b[3];
    space2body(b, q, x);
    double       nrm = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    double       dn  = 1.0 / nrm;
    const double r_[3] = {b[0] * dn, b[1] * dn, b[2] * dn};
    double       ph[3];
    vcross(ph, e3, r_);
    const double dot = e3[0] * r_[0] + e3[1] * r_[1] + e3[2] * r_[2];
    dn               = sqrt(ph[0] * ph[0] + ph[1] * ph[1] + ph[2] * ph[2]);
    const double tol = 10.0 * 2.220446049250313080847e-16;
    if (((dn > tol) || (-tol > dn)) && less_than_mp(fabs(dot), 1.0)) {
        dn = 1.0 / dn;
        for (int d = 0; d < 3; d++) ph[d] *= dn;
        *theta = acos(dot);
        double th[3];
        vcross(th, ph, r_);
        dn = 1.0 / sqrt(th[0] * th[0] + th[1] * th[1] + th[2] * th[2]);
        for (int d = 0; d < 3; d++) th[d] *= dn;
        body2space(nr, q, r_);
        body2space(nth, q, th);
        body2space(nph, q, ph);
    } else {
        nrm = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        dn  = 1.0 / nrm;
        for (int d = 0; d < 3; d++) {
            nr[d]  = x[d] * dn;
            nth[d] = nph[d] = 0.0;
        }
        *theta = (dot > 0.0) ? 0.0 : 3.14159265358979323846;
    }
}

// Make_particle_momentum_factor (operate_surface.cxx:12-153): mass moments
// of phi and of the surface weight phi_s = (1 - phi) DPhi_compact_sin_norm,
// and the slip momentum the shell carries against the fluid u.
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_slip_moments(PParams pp, DevParticle *__restrict__ P,
                                                        const T *__restrict__ u0, const T *__restrict__ u1,
                                                        const T *__restrict__ u2) {
    const int         n    = blockIdx.x;
    DevParticle      &p    = P[n];
    const DevSpecies &s    = pp.species[p.spec];
    const bool        slip = s.propulsion == JP_SLIP;  // mass moments for every particle, slip sums for squirmers
    const double xp[3] = {p.x[0], p.x[1], p.x[2]};
    const double q[4]  = {p.q[0], p.q[1], p.q[2], p.q[3]};
    const PCell  c     = particle_cell(pp, xp);
    const T     *U[3]  = {u0, u1, u2};
    Neu          acc[32];  // M0, M1[3], M2[9], SM0, SM1[3], SM2[9], dv[3], dw[3]
    for (int m = threadIdx.x; m < pp.np_domain; m += PT_NT) {
        const PPoint pt  = sek_point(pp, c, xp, pp.sek[m]);
        const double r   = pt.dist;
        const double ph  = prof_phi(pp, r);
        const double *rv = pt.r;
        acc[0].add(ph);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            acc[1 + d].add(ph * rv[d]);
#pragma unroll
            for (int l = 0; l < 3; l++) acc[4 + 3 * d + l].add((d == l ? ph * r * r : 0.0) - ph * rv[d] * rv[l]);
        }
        if (slip && less_than_mp(fabs(r - pp.radius), pp.hxi)) {
            const double phs = (1.0 - ph) * dphi_sin(pp, r);
            double       nr[3], nth[3], nph[3], th;
            squirmer_coord(rv, q, s.janus_axis, nr, nth, nph, &th);
            const double mag = s.slip_vel * (sin(th) + s.slip_mode_half * sin(2.0 * th));
            const double az  = s.c1 * sin(th) + s.c2 * 1.5 * sin(2.0 * th);
            double       us[3], rxus[3];
            for (int d = 0; d < 3; d++) us[d] = nth[d] * mag + nph[d] * az - (double)U[d][pt.idx];
            vcross(rxus, rv, us);
            acc[13].add(phs);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                acc[14 + d].add(phs * rv[d]);
#pragma unroll
                for (int l = 0; l < 3; l++) acc[17 + 3 * d + l].add((d == l ? phs * r * r : 0.0) - phs * rv[d] * rv[l]);
                acc[26 + d].add(phs * us[d]);
                acc[29 + d].add(phs * rxus[d]);
            }
        }
    }
    double v[32];
    for (int k = 0; k < 32; k++) v[k] = acc[k].val();
    block_sum<32>(v);
    if (threadIdx.x == 0) {
        p.mass  = v[0];
        p.smass = v[13];
        for (int d = 0; d < 3; d++) {
            p.mass_center[d] = v[1 + d];
            p.smc[d]         = v[14 + d];
            p.sdv[d]         = v[26 + d];
            p.sdw[d]         = v[29 + d];
        }
        for (int k = 0; k < 9; k++) {
            p.inertia[k]  = v[4 + k];
            p.sinertia[k] = v[17 + k];
        }
    }
}

// slip_droplet (operate_surface.cxx:155-198): the droplet counter-flow
// (delta_v, delta_w) that cancels the shell's linear and angular momentum.
__device__ __forceinline__ void slip_droplet(const DevParticle &p, double *vp, double *wp, double *dv, double *dw) {
    double LL[3][3], dP[3], dL[3], L0[3], mc[3];
    for (int d = 0; d < 3; d++) {
        vp[d] = p.v_slip[d];
        wp[d] = p.om_slip[d];
        mc[d] = p.mass_center[d];
    }
    const double imass = 1.0 / p.mass;
    double       wxr[3], rxv[3];
    vcross(wxr, wp, p.smc);
    vcross(rxv, p.smc, vp);
    for (int d = 0; d < 3; d++) {
        dP[d] = -(p.smass * vp[d] + wxr[d] + p.sdv[d]);
        dL[d] = -(rxv[d] + p.sinertia[3 * d] * wp[0] + p.sinertia[3 * d + 1] * wp[1] + p.sinertia[3 * d + 2] * wp[2] +
                  p.sdw[d]);
    }
    double t[3];
    vcross(t, mc, dP);
    const double mc2 = -(mc[0] * mc[0] + mc[1] * mc[1] + mc[2] * mc[2]) * imass;
    for (int d = 0; d < 3; d++) {
        L0[d] = dL[d] - imass * t[d];
        for (int l = 0; l < 3; l++) LL[d][l] = (d == l ? mc2 : 0.0) + (p.inertia[3 * d + l] + imass * mc[d] * mc[l]);
    }
    // M_inv (lad3.h:202-233)
    const double(&A)[3][3] = LL;
    double idet = -A[0][2] * A[1][1] * A[2][0] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
                  A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] + A[0][0] * A[1][1] * A[2][2];
    idet = 1.0 / idet;
    double B[3][3];
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * idet;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * idet;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * idet;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * idet;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * idet;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * idet;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * idet;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * idet;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * idet;
    for (int i = 0; i < 3; i++) dw[i] = B[i][0] * L0[0] + B[i][1] * L0[1] + B[i][2] * L0[2];
    vcross(t, dw, mc);
    for (int d = 0; d < 3; d++) dv[d] = imass * (dP[d] - t[d]);
}

// Make_force_u_slip_particle (operate_surface.cxx:200-363): scatter the
// droplet counter-flow phi (dv + dw x r) and the shell slip force
// phi_s (v + w x r + u_s - u) into up; f_slip, torque_slip reset.
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_slip_scatter(PParams pp, DevParticle *__restrict__ P,
                                                        const T *__restrict__ u0, const T *__restrict__ u1,
                                                        const T *__restrict__ u2, T *__restrict__ up0,
                                                        T *__restrict__ up1, T *__restrict__ up2,
                                                        const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int         n = blockIdx.x;
    DevParticle      &p = P[n];
    const DevSpecies &s = pp.species[p.spec];
    __shared__ double sh[12];
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) {
            p.fs[d] = 0.0;
            p.ts[d] = 0.0;
        }
        if (s.propulsion == JP_SLIP) slip_droplet(p, sh, sh + 3, sh + 6, sh + 9);
    }
    if (s.propulsion != JP_SLIP) return;
    __syncthreads();
    const double vp[3] = {sh[0], sh[1], sh[2]}, wp[3] = {sh[3], sh[4], sh[5]};
    const double dv[3] = {sh[6], sh[7], sh[8]}, dw[3] = {sh[9], sh[10], sh[11]};
    const double xp[3] = {p.x[0], p.x[1], p.x[2]};
    const double q[4]  = {p.q[0], p.q[1], p.q[2], p.q[3]};
    const PCell  c     = particle_cell(pp, xp);
    const T     *U[3]  = {u0, u1, u2};
    T           *UP[3] = {up0, up1, up2};
    for (int m = threadIdx.x; m < pp.np_domain; m += PT_NT) {
        const PPoint  pt = sek_point(pp, c, xp, pp.sek[m]);
        const double  r  = pt.dist;
        const double  ph = prof_phi(pp, r);
        const double *rv = pt.r;
        double        vr[3], f[3];
        angular2v(pp.rotation, dw, rv, vr);
        for (int d = 0; d < 3; d++) f[d] = ph * (dv[d] + vr[d]);
        const double phs = (1.0 - ph) * dphi_sin(pp, r);
        if (less_than_mp(fabs(r - pp.radius), pp.hxi)) {
            double vrot[3], nr[3], nth[3], nph[3], th;
            angular2v(pp.rotation, wp, rv, vrot);
            squirmer_coord(rv, q, s.janus_axis, nr, nth, nph, &th);
            const double vs  = s.slip_vel * (sin(th) + s.slip_mode_half * sin(2.0 * th));
            const double vs2 = s.c1 * sin(th) + s.c2 * 1.5 * sin(2.0 * th);
            for (int d = 0; d < 3; d++)
                f[d] += phs * (vp[d] + vrot[d] + nth[d] * vs + nph[d] * vs2 - (double)U[d][pt.idx]);
        }
        for (int d = 0; d < 3; d++)
            if (f[d] != 0.0) unsafeAtomicAdd(&UP[d][pt.idx], (T)f[d]);
    }
}

// Calc_f_slip_correct_precision (md_force.cxx:325-382): momentum the
// projected slip field carries over each particle's profile.
template <typename T>
__global__ __launch_bounds__(PT_NT) void k_f_slip(PParams pp, DevParticle *__restrict__ P, const T *__restrict__ s0,
                                                  const T *__restrict__ s1, const T *__restrict__ s2, double coef,
                                                  const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int     n     = blockIdx.x;
    const double  xp[3] = {P[n].x[0], P[n].x[1], P[n].x[2]};
    const PCell   c     = particle_cell(pp, xp);
    Neu           acc[6];
    for (int m = threadIdx.x; m < pp.np_domain; m += PT_NT) {
        const PPoint pt = sek_point(pp, c, xp, pp.sek[m]);
        const double ph = prof_phi(pp, pt.dist);
        if (ph == 0.0) continue;
        const double f[3] = {(double)s0[pt.idx] * ph, (double)s1[pt.idx] * ph, (double)s2[pt.idx] * ph};
        acc[0].add(f[0]);
        acc[1].add(f[1]);
        acc[2].add(f[2]);
        acc[3].add(pt.r[1] * f[2] - pt.r[2] * f[1]);
        acc[4].add(pt.r[2] * f[0] - pt.r[0] * f[2]);
        acc[5].add(pt.r[0] * f[1] - pt.r[1] * f[0]);
    }
    double v[6];
    for (int k = 0; k < 6; k++) v[k] = acc[k].val();
    block_sum<6>(v);
    if (threadIdx.x == 0)
        for (int d = 0; d < 3; d++) {
            P[n].fs[d] = coef * v[d];
            P[n].ts[d] = coef * v[3 + d];
        }
}

// MD_solver_velocity_slip_Euler / _AB2 (particle_solver.cxx:144-207):
// the velocity update with the slip force, forces kept for the iteration.
__global__ void k_velocity_slip(PParams pp, DevParticle *__restrict__ P, int ab2, double h,
                                const int *__restrict__ skip) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pp.np || (skip && *skip)) return;
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
                p.v[d] += dmy * (p.fh[d] + p.fs[d] + p.fr[d] + sf[d]);
                p.om[d] += dmy_rot * (p.th[d] + p.ts[d] + p.tr[d] + st[d]);
            } else {
                p.v[d] += dmy * (2.0 * (p.fh[d] + p.fs[d] + sf[d]) + p.fr[d] + p.fr_prev[d]);
                p.om[d] += dmy_rot * (2.0 * (p.th[d] + p.ts[d] + st[d]) + p.tr[d] + p.tr_prev[d]);
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
}

// reset_Forces (particle_solver.cxx:11-37)
__global__ void k_reset_forces(PParams pp, DevParticle *__restrict__ P) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pp.np) return;
    DevParticle &p = P[n];
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

// Update_slip_particle_velocity (operate_surface.h:101-121); with `restore`
// the velocities go back to the values before the iteration (reset_iter).
__global__ void k_slip_update(PParams pp, DevParticle *__restrict__ P, int iter, int restore,
                              const int *__restrict__ skip) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= pp.np || (skip && *skip)) return;
    DevParticle &p = P[n];
    for (int d = 0; d < 3; d++) {
        if (iter == 0) {
            p.v_slip[d]  = p.v[d];
            p.om_slip[d] = p.om[d];
        } else {
            p.v_slip[d]  = 0.3 * p.v_slip[d] + 0.7 * p.v[d];
            p.om_slip[d] = 0.3 * p.om_slip[d] + 0.7 * p.om[d];
        }
        if (restore) {
            p.v[d]  = p.v_old[d];
            p.om[d] = p.om_old[d];
        }
    }
}

// Slip_particle_convergence (operate_surface.h:72-100): max over slip
// particles of |v_slip - v|^2/|v|^2 + |w_slip - w|^2/|w|^2, square-rooted.
// With a loop decided on the device (flag non-null): nothing once flag[0] is
// set; otherwise the iteration is counted and flag[0] set when eps < tol or
// the count reaches max_iter (the exit test of sp_3d_ns.cxx:153-158).
__global__ __launch_bounds__(PT_NT) void k_slip_conv(PParams pp, const DevParticle *__restrict__ P,
                                                     double *__restrict__ eps_out, int *__restrict__ flag,
                                                     unsigned long long *__restrict__ stat, double tol,
                                                     int max_iter) {
    if (flag && flag[0]) return;
    __shared__ double red[PT_NT / 64];
    double            e = 0.0;
    for (int n = threadIdx.x; n < pp.np; n += PT_NT) {
        const DevParticle &p = P[n];
        if (pp.species[p.spec].propulsion != JP_SLIP) continue;
        double dv = 0, dw = 0, nv = 0, nw = 0;
        for (int d = 0; d < 3; d++) {
            nv += p.v[d] * p.v[d];
            nw += p.om[d] * p.om[d];
            dv += (p.v_slip[d] - p.v[d]) * (p.v_slip[d] - p.v[d]);
            dw += (p.om_slip[d] - p.om[d]) * (p.om_slip[d] - p.om[d]);
        }
        const double tol = 10.0 * 2.220446049250313080847e-16;
        if (nv > tol) dv /= nv;
        if (nw > tol) dw /= nw;
        e = fmax(e, dv + dw);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e = fmax(e, __shfl_down(e, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < PT_NT / 64; w++) e = fmax(e, red[w]);
        const double eps = sqrt(e);
        eps_out[0]       = eps;
        if (flag) {
            const int it = flag[1] + 1;
            flag[1]      = it;
            stat[0] += 1;
            if (eps < tol || it == max_iter) {
                flag[0] = 1;
                stat[1] += 1;
            }
        }
    }
}

__global__ void k_slip_loop_begin(int *__restrict__ flag) {
    if (threadIdx.x == 0) {
        flag[0] = 0;
        flag[1] = 0;
    }
}

// ------------------------------------------------------------ launchers
static inline unsigned pblocks(int n) { return (unsigned)((n + 63) / 64); }

// Workgroup size of the phi scatter and the fused particle step: 1024 lanes
// while the particles' workgroups (16 wavefronts, one per CU at the step's
// register count) all fit on the chip at once, 256 beyond.
constexpr int PT_WIDE = 1024;
static inline bool pt_wide(const PParams &pp) { return pp.np <= 256; }

template <typename T, int NT, bool OBL>
static void launch_phi_scatter_v(const PParams &pp, DevParticle *P, T *phisum, const PCache<T> &pc, int scatter,
                                 hipStream_t s, int move, int ab2, double dt, double hdt) {
    if (move)
        hipLaunchKernelGGL((k_phi_scatter<T, true, true, NT, OBL>), dim3(pp.np), dim3(NT), 0, s, pp, P, phisum, pc,
                           ab2, dt, hdt);
    else if (scatter)
        hipLaunchKernelGGL((k_phi_scatter<T, true, false, NT, OBL>), dim3(pp.np), dim3(NT), 0, s, pp, P, phisum, pc,
                           ab2, dt, hdt);
    else
        hipLaunchKernelGGL((k_phi_scatter<T, false, false, NT, OBL>), dim3(pp.np), dim3(NT), 0, s, pp, P, phisum, pc,
                           ab2, dt, hdt);
}
template <typename T, int NT>
static void launch_phi_scatter(const PParams &pp, DevParticle *P, T *phisum, const PCache<T> &pc, int scatter,
                               hipStream_t s, int move, int ab2, double dt, double hdt) {
    if (pp.obl)
        launch_phi_scatter_v<T, NT, true>(pp, P, phisum, pc, scatter, s, move, ab2, dt, hdt);
    else
        launch_phi_scatter_v<T, NT, false>(pp, P, phisum, pc, scatter, s, move, ab2, dt, hdt);
}
template <typename T>
hipError_t phi_scatter(const PParams &pp, DevParticle *P, T *phisum, const PCache<T> &pc, int scatter, hipStream_t s,
                       int move, int ab2, double dt, double hdt) {
    if (pp.np > 0) {
        if (pt_wide(pp))
            launch_phi_scatter<T, PT_WIDE>(pp, P, phisum, pc, scatter, s, move, ab2, dt, hdt);
        else
            launch_phi_scatter<T, PT_NT>(pp, P, phisum, pc, scatter, s, move, ab2, dt, hdt);
    }
    return hipGetLastError();
}
template <typename T>
hipError_t f_hydro(const PParams &pp, DevParticle *P, T *const u[3], const T *phisum, const PCache<T> &pc,
                   double coef, hipStream_t s) {
    if (pp.np > 0) {
        if (pp.obl)
            hipLaunchKernelGGL((k_f_hydro<T, true>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, u[0], u[1], u[2], phisum,
                               pc, coef);
        else
            hipLaunchKernelGGL((k_f_hydro<T, false>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, u[0], u[1], u[2], phisum,
                               pc, coef);
    }
    return hipGetLastError();
}
template <typename T>
hipError_t up_scatter(const PParams &pp, const DevParticle *P, const T *phisum, const PCache<T> &pc,
                      T *const up[3], hipStream_t s) {
    if (pp.np > 0) {
        if (pp.obl)
            hipLaunchKernelGGL((k_up_scatter<T, true>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, phisum, pc, up[0],
                               up[1], up[2]);
        else
            hipLaunchKernelGGL((k_up_scatter<T, false>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, phisum, pc, up[0],
                               up[1], up[2]);
    }
    return hipGetLastError();
}
hipError_t position(const PParams &pp, DevParticle *P, int ab2, double dt, double hdt, hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_position, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, ab2, dt, hdt);
    return hipGetLastError();
}
hipError_t forces(const PParams &pp, DevParticle *P, hipStream_t s) {
    if (pp.np > 0 && pp.bid)
        hipLaunchKernelGGL((k_forces<false, true>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P);
    else if (pp.np > 0 && pp.obl)
        hipLaunchKernelGGL(k_forces<true>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P);
    else if (pp.np > 0)
        hipLaunchKernelGGL(k_forces<false>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P);
    return hipGetLastError();
}
hipError_t velocity(const PParams &pp, DevParticle *P, int ab2, double h, hipStream_t s, int noise, uint64_t seed,
                    uint32_t stream, const int *d_ts, double niv, double nio) {
    const Noise nz{noise, seed, stream, d_ts, niv, nio};
    if (pp.np > 0) hipLaunchKernelGGL(k_velocity, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, ab2, h, nz);
    return hipGetLastError();
}
hipError_t thermostat(const PParams &pp, DevParticle *P, uint64_t seed, uint32_t stream, const int *d_ts, double niv,
                      double nio, hipStream_t s) {
    const Noise nz{1, seed, stream, d_ts, niv, nio};
    if (pp.np > 0) hipLaunchKernelGGL(k_thermostat, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, nz);
    return hipGetLastError();
}

// ---------------------------------------------------------- rigid bodies
__device__ __forceinline__ void cross3(double *c, const double *a, const double *b) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// solver_Rigid_Position (rigid.h:333-392) with update_Orientation (271-290)
// and update_Particle_Configuration (159-178): lane 0 moves the body's
// centre and turns its first bead's quaternion (MD_solver_orientation_*,
// from that bead's omega / omega_old); then the lanes place the beads.
__global__ __launch_bounds__(64) void k_rigid_position(PParams pp, DevParticle *__restrict__ P, RigidArgs ra, int ab2,
                                                       double dt, double hdt) {
    DevBody &B = ra.B[blockIdx.x];
    __shared__ double sq[4], sqo[4], sx[3], sv[3], sw[3];
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) {
            B.xGp[d]             = B.xG[d];
            const double delta_x = ab2 ? hdt * (3.0 * B.vG[d] - B.vGo[d]) : dt * B.vG[d];
            B.xG[d] += delta_x;
            B.xGn[d] += delta_x;
            B.xG[d] = fmod(B.xG[d] + pp.L[d], pp.L[d]);
        }
        DevParticle &f = P[B.first];
        double       dq[4];
        if (!ab2) {
            for (int k = 0; k < 4; k++) f.q_old[k] = f.q[k];
            qdot_space(dq, f.q, f.om);
            for (int k = 0; k < 4; k++) f.q[k] += dq[k] * dt;
        } else {
            double dqo[4];
            qdot_space(dq, f.q, f.om);
            qdot_space(dqo, f.q_old, f.om_old);
            for (int k = 0; k < 4; k++) f.q_old[k] = f.q[k];
            for (int k = 0; k < 4; k++) f.q[k] += dq[k] * (3.0 * hdt);
            for (int k = 0; k < 4; k++) f.q[k] += dqo[k] * (-hdt);
        }
        qtn_normalize(f.q);
        for (int k = 0; k < 4; k++) {
            sq[k]  = f.q[k];
            sqo[k] = f.q_old[k];
        }
        for (int d = 0; d < 3; d++) {
            sx[d] = B.xG[d];
            sv[d] = B.vG[d];
            sw[d] = B.wG[d];
        }
    }
    __syncthreads();
    for (int n = B.first + threadIdx.x; n < B.last; n += 64) {
        DevParticle &p = P[n];
        if (n != B.first)
            for (int k = 0; k < 4; k++) {
                p.q_old[k] = sqo[k];
                p.q[k]     = sq[k];
            }
        double gr[3], wxr[3];
        body2space(gr, sq, ra.GRb + 3 * (long long)n);
        cross3(wxr, sw, gr);
        for (int d = 0; d < 3; d++) {
            ra.GR[3 * (long long)n + d] = gr[d];
            p.x_prev[d]                 = p.x[d];
            p.x[d]                      = fmod(sx[d] + gr[d] + pp.L[d], pp.L[d]);
            p.om_old[d]                 = p.om[d];
            p.v_old[d]                  = p.v[d];
            p.om[d]                     = sw[d];
            p.v[d]                      = sv[d] + wxr[d];
            pp.xs[4 * n + d]            = p.x[d];
        }
    }
}

// calc_Rigid_VOGs (rigid.h:467-559) after Force: lane 0 sums its beads'
// hydrodynamic force and torque about the centre (forceGs, torqueGs: GR x
// f_hydro + torque_hydro per bead, md_force.cxx:467-505) and LJ force and
// torque (forceGrs, torqueGrs, md_force.cxx:134-153), adds gravity on the
// body's volume (md_force.cxx:316-323), then integrates the centre velocity
// and Euler's equations in the body frame (particle_rotation_solver.h:
// 106-184); the lanes then give the beads the body's motion
// (set_Particle_Velocities, rigid.h:218-229) and roll their forces
// (reset_Forces, particle_solver.cxx:11-37).
__global__ __launch_bounds__(64) void k_rigid_velocity(PParams pp, DevParticle *__restrict__ P, RigidArgs ra,
                                                       int ab2, double dt, double hdt) {
    DevBody &B = ra.B[blockIdx.x];
    __shared__ double sv[3], sw[3];
    if (threadIdx.x == 0) {
        double fG[3] = {0, 0, 0}, tG[3] = {0, 0, 0}, fGr[3] = {0, 0, 0}, tGr[3] = {0, 0, 0};
        for (int n = B.first; n < B.last; n++) {
            const DevParticle &p  = P[n];
            const double      *gr = ra.GR + 3 * (long long)n;
            double             gxf[3], gxr[3], fl[3];
            cross3(gxf, gr, p.fh);
            for (int d = 0; d < 3; d++) {
                fG[d] += p.fh[d];
                tG[d] += gxf[d] + p.th[d];
                fGr[d] += p.fr[d];
                // the reference adds GR x f_i to the body of the partner of lower
                // index too, whose bead takes -f_i: per bead, the forces from
                // partners of higher index enter with their sign flipped
                fl[d] = ra.compat ? p.fr[d] - 2.0 * pp.flo[3 * (long long)n + d] : p.fr[d];
            }
            cross3(gxr, gr, fl);
            for (int d = 0; d < 3; d++) tGr[d] += gxr[d];
        }
        const DevSpecies &sp = pp.species[B.spec];
        if (pp.g != 0.0) fGr[pp.g_dir] -= pp.g * pp.rho * (B.M / (sp.mass_ratio * pp.rho)) * (sp.mass_ratio - 1.0);
        for (int d = 0; d < 3; d++) {
            B.vGo[d] = B.vG[d];
            B.wGo[d] = B.wG[d];
        }
        const double *q = P[B.first].q, *qo = P[B.first].q_old;
        double        wb[3], tb[3], tbo[3], tq[3], nw[3];
        if (!ab2) {
            if (B.free_)
                for (int d = 0; d < 3; d++) B.vG[d] += (dt / B.M) * (fG[d] + fGr[d]);
            for (int d = 0; d < 3; d++) tq[d] = tG[d] + tGr[d];
            space2body(wb, q, B.wG);
            space2body(tb, q, tq);
            for (int d = 0; d < 3; d++) nw[d] = wb[d];
            if (B.free_) {
                nw[0] += (dt / B.Ib[0]) * (tb[0] + (B.Ib[1] - B.Ib[2]) * wb[1] * wb[2]);
                nw[1] += (dt / B.Ib[1]) * (tb[1] + (B.Ib[2] - B.Ib[0]) * wb[2] * wb[0]);
                nw[2] += (dt / B.Ib[2]) * (tb[2] + (B.Ib[0] - B.Ib[1]) * wb[0] * wb[1]);
            }
        } else {
            if (B.free_)
                for (int d = 0; d < 3; d++) B.vG[d] += (hdt / B.M) * (2.0 * fG[d] + fGr[d] + B.fGrp[d]);
            for (int d = 0; d < 3; d++) tq[d] = 2.0 * tG[d] + tGr[d];
            space2body(wb, q, B.wG);
            space2body(tb, q, tq);
            space2body(tbo, qo, B.tGrp);
            for (int d = 0; d < 3; d++) nw[d] = wb[d];
            if (B.free_) {
                nw[0] += (hdt / B.Ib[0]) * ((tb[0] + tbo[0]) + 2.0 * (B.Ib[1] - B.Ib[2]) * wb[1] * wb[2]);
                nw[1] += (hdt / B.Ib[1]) * ((tb[1] + tbo[1]) + 2.0 * (B.Ib[2] - B.Ib[0]) * wb[2] * wb[0]);
                nw[2] += (hdt / B.Ib[2]) * ((tb[2] + tbo[2]) + 2.0 * (B.Ib[0] - B.Ib[1]) * wb[0] * wb[1]);
            }
        }
        body2space(B.wG, q, nw);
        for (int d = 0; d < 3; d++) {
            B.fGp[d]  = fG[d];
            B.tGp[d]  = tG[d];
            B.fGrp[d] = fGr[d];
            B.tGrp[d] = tGr[d];
            sv[d]     = B.vG[d];
            sw[d]     = B.wG[d];
        }
    }
    __syncthreads();
    for (int n = B.first + threadIdx.x; n < B.last; n += 64) {
        DevParticle &p = P[n];
        double       wxr[3];
        cross3(wxr, sw, ra.GR + 3 * (long long)n);
        for (int d = 0; d < 3; d++) {
            p.om[d]      = sw[d];
            p.v[d]       = sv[d] + wxr[d];
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
}

hipError_t rigid_position(const PParams &pp, DevParticle *P, const RigidArgs &ra, int ab2, double dt, double hdt,
                          hipStream_t s) {
    if (ra.nb > 0) hipLaunchKernelGGL(k_rigid_position, dim3(ra.nb), dim3(64), 0, s, pp, P, ra, ab2, dt, hdt);
    return hipGetLastError();
}
hipError_t rigid_velocity(const PParams &pp, DevParticle *P, const RigidArgs &ra, int ab2, double dt, double hdt,
                          hipStream_t s) {
    if (ra.nb > 0) hipLaunchKernelGGL(k_rigid_velocity, dim3(ra.nb), dim3(64), 0, s, pp, P, ra, ab2, dt, hdt);
    return hipGetLastError();
}

// jikan.ts += 1 on the device (the step counter the noise stream is keyed by)
__global__ void k_tick(int *__restrict__ d_ts, Grid bg) {
    if (batch_skip(bg)) return;
    if (threadIdx.x == 0 && blockIdx.x == 0) *brun(d_ts, bg) += 1;
}
hipError_t tick(int *d_ts, hipStream_t s, const Grid *batch) {
    const Grid bg = batch ? *batch : Grid{};
    hipLaunchKernelGGL(k_tick, dim3(1, 1, batch_z(bg)), dim3(64), 0, s, d_ts, bg);
    return hipGetLastError();
}

template <typename T>
hipError_t particle_step(const PParams &pp, DevParticle *P, T *const u[3], const T *phisum, const PCache<T> &pc,
                         double coef, int ab2, double h, int noise, uint64_t seed, uint32_t stream, const int *d_ts,
                         double niv, double nio, T *const up[3], hipStream_t s) {
    const Noise nz{noise, seed, stream, d_ts, niv, nio};
    if (pp.np > 0 && pt_wide(pp))
        hipLaunchKernelGGL((k_particle_step<T, PT_WIDE>), dim3(pp.np), dim3(PT_WIDE), 0, s, pp, P, u[0], u[1], u[2],
                           phisum, pc, coef, ab2, h, nz, up[0], up[1], up[2]);
    else if (pp.np > 0)
        hipLaunchKernelGGL((k_particle_step<T, PT_NT>), dim3(pp.np), dim3(PT_NT), 0, s, pp, P, u[0], u[1], u[2],
                           phisum, pc, coef, ab2, h, nz, up[0], up[1], up[2]);
    return hipGetLastError();
}
template <typename T>
hipError_t slip_moments(const PParams &pp, DevParticle *P, T *const u[3], hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_slip_moments<T>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P, u[0], u[1], u[2]);
    return hipGetLastError();
}
template <typename T>
hipError_t slip_scatter(const PParams &pp, DevParticle *P, T *const u[3], T *const up[3], hipStream_t s,
                        const int *skip) {
    if (pp.np > 0)
        hipLaunchKernelGGL(k_slip_scatter<T>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P, u[0], u[1], u[2], up[0], up[1],
                           up[2], skip);
    return hipGetLastError();
}
template <typename T>
hipError_t f_slip(const PParams &pp, DevParticle *P, T *const sf[3], double coef, hipStream_t s, const int *skip) {
    if (pp.np > 0)
        hipLaunchKernelGGL(k_f_slip<T>, dim3(pp.np), dim3(PT_NT), 0, s, pp, P, sf[0], sf[1], sf[2], coef, skip);
    return hipGetLastError();
}
hipError_t velocity_slip(const PParams &pp, DevParticle *P, int ab2, double h, hipStream_t s, const int *skip) {
    if (pp.np > 0) hipLaunchKernelGGL(k_velocity_slip, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, ab2, h, skip);
    return hipGetLastError();
}
hipError_t reset_forces(const PParams &pp, DevParticle *P, hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_reset_forces, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P);
    return hipGetLastError();
}
hipError_t slip_update(const PParams &pp, DevParticle *P, int iter, int restore, hipStream_t s, const int *skip) {
    if (pp.np > 0)
        hipLaunchKernelGGL(k_slip_update, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, iter, restore, skip);
    return hipGetLastError();
}
hipError_t slip_conv(const PParams &pp, const DevParticle *P, double *eps, hipStream_t s, const SlipLoop *loop) {
    hipLaunchKernelGGL(k_slip_conv, dim3(1), dim3(PT_NT), 0, s, pp, P, eps, loop ? loop->flag : nullptr,
                       loop ? loop->stat : nullptr, loop ? loop->tol : 0.0, loop ? loop->max_iter : 0);
    return hipGetLastError();
}
hipError_t slip_loop_begin(const SlipLoop &loop, hipStream_t s) {
    hipLaunchKernelGGL(k_slip_loop_begin, dim3(1), dim3(64), 0, s, loop.flag);
    return hipGetLastError();
}

// Per-step particle record: after the step that made ts = *d_ts, slot
// ts - base - 1 receives (x, v) of every particle; slots past nslots are
// not written.
__device__ __forceinline__ void record_one(const PParams &pp, const DevParticle *__restrict__ P,
                                           const int *__restrict__ d_ts, int base, int nslots,
                                           double *__restrict__ out, int n) {
    const int slot = *d_ts - base - 1;
    if (n >= pp.np || slot < 0 || slot >= nslots) return;
    double *o = out + ((size_t)slot * pp.np + n) * 6;
    for (int d = 0; d < 3; d++) {
        o[d]     = P[n].x[d];
        o[3 + d] = P[n].v[d];
    }
}
__global__ void k_record(PParams pp, const DevParticle *__restrict__ P, const int *__restrict__ d_ts, int base,
                         int nslots, double *__restrict__ out) {
    record_one(pp, P, d_ts, base, nslots, out, blockIdx.x * blockDim.x + threadIdx.x);
}
hipError_t record(const PParams &pp, const DevParticle *P, const int *d_ts, int base, int nslots, double *out,
                  hipStream_t s) {
    if (pp.np > 0) hipLaunchKernelGGL(k_record, dim3(pblocks(pp.np)), dim3(64), 0, s, pp, P, d_ts, base, nslots, out);
    return hipGetLastError();
}

// One Output_udf frame (output.cxx:268-375): when the step counter ts has
// reached ts0 + k*every (k < nframes), frame k = {ts, then per particle R,
// R_raw, v, q (q0..q3), omega, f_hydro, torque_hydro, f_r, torque_r (0),
// f_slip, torque_slip}, the forces and torques being the previous step's, as
// the reference writes them.
__device__ __forceinline__ void frame_one(const PParams &pp, const DevParticle *__restrict__ P,
                                          const int *__restrict__ d_ts, int ts0, int every, int nframes,
                                          double *__restrict__ out, int n) {
    const int ts = *d_ts, dt = ts - ts0;
    if (dt < 0 || dt % every != 0 || dt / every >= nframes) return;
    double *f = out + (size_t)(dt / every) * (1 + (size_t)FRAME_WORDS * pp.np);
    if (n == 0) f[0] = (double)ts;
    if (n >= pp.np) return;
    const DevParticle &q = P[n];
    double            *o = f + 1 + (size_t)FRAME_WORDS * n;
    const double *src[10] = {q.x, q.x_nopbc, q.v, q.om, q.fh_prev, q.th_prev, q.fr_prev, nullptr, q.fs_prev, q.ts_prev};
    int           k       = 0;
    for (int s = 0; s < 10; s++) {
        for (int d = 0; d < 3; d++) o[k + d] = src[s] ? src[s][d] : 0.0;
        k += 3;
        if (s == 2) {  // q after v
            for (int d = 0; d < 4; d++) o[k + d] = q.q[d];
            k += 4;
        }
    }
}
__global__ void k_frame(PParams pp, const DevParticle *__restrict__ P, const int *__restrict__ d_ts, int ts0,
                        int every, int nframes, double *__restrict__ out) {
    frame_one(pp, P, d_ts, ts0, every, nframes, out, blockIdx.x * blockDim.x + threadIdx.x);
}
// the per-step record and frames of every stepped run of a batch (grid z = slot)
template <typename T>
__global__ void k_records_batch(const SlotPart<T> *__restrict__ sp, const int *__restrict__ active) {
    if (!active[blockIdx.z]) return;
    const SlotPart<T> &a = sp[blockIdx.z];
    for (int n = threadIdx.x; n < (a.pp.np > 0 ? a.pp.np : 1); n += blockDim.x) {
        if (a.traj) record_one(a.pp, a.P, a.nz.d_ts, a.traj_base, a.traj_steps, a.traj, n);
        if (a.frames) frame_one(a.pp, a.P, a.nz.d_ts, a.frame_ts0, a.frame_every, a.frame_n, a.frames, n);
    }
}
hipError_t frame(const PParams &pp, const DevParticle *P, const int *d_ts, int ts0, int every, int nframes, double *out,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_frame, dim3(pblocks(pp.np > 0 ? pp.np : 1)), dim3(64), 0, s, pp, P, d_ts, ts0, every, nframes,
                       out);
    return hipGetLastError();
}

template <typename T>
hipError_t phi_scatter_batch(const SlotPart<T> *sp, const int *active, int nslots, int np_cap, double dt, double hdt,
                             hipStream_t s) {
    if (np_cap < 1 || np_cap > 256) return hipErrorInvalidValue;  // the 1024-lane form of few particles
    hipLaunchKernelGGL((k_phi_scatter_batch<T, PT_WIDE>), dim3(np_cap, 1, nslots), dim3(PT_WIDE), 0, s, sp, active, dt,
                       hdt);
    return hipGetLastError();
}
template <typename T>
hipError_t particle_step_batch(const SlotPart<T> *sp, const int *active, int nslots, int np_cap, double coef, double h,
                               hipStream_t s) {
    if (np_cap < 1 || np_cap > 256) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_particle_step_batch<T, PT_WIDE>), dim3(np_cap, 1, nslots), dim3(PT_WIDE), 0, s, sp, active,
                       coef, h);
    return hipGetLastError();
}
template <typename T>
hipError_t records_batch(const SlotPart<T> *sp, const int *active, int nslots, hipStream_t s) {
    hipLaunchKernelGGL(k_records_batch<T>, dim3(1, 1, nslots), dim3(64), 0, s, sp, active);
    return hipGetLastError();
}

#define KH_INST(T)                                                                                               \
    template hipError_t phi_scatter_batch<T>(const SlotPart<T> *, const int *, int, int, double, double,          \
                                             hipStream_t);                                                       \
    template hipError_t particle_step_batch<T>(const SlotPart<T> *, const int *, int, int, double, double,        \
                                               hipStream_t);                                                     \
    template hipError_t records_batch<T>(const SlotPart<T> *, const int *, int, hipStream_t);                     \
    template hipError_t phi_scatter<T>(const PParams &, DevParticle *, T *, const PCache<T> &, int, hipStream_t, int, \
                                       int, double, double);                                                       \
    template hipError_t f_hydro<T>(const PParams &, DevParticle *, T *const[3], const T *, const PCache<T> &,       \
                                   double, hipStream_t);                                                           \
    template hipError_t up_scatter<T>(const PParams &, const DevParticle *, const T *, const PCache<T> &,           \
                                      T *const[3], hipStream_t);                                                   \
    template hipError_t slip_moments<T>(const PParams &, DevParticle *, T *const[3], hipStream_t);                  \
    template hipError_t particle_step<T>(const PParams &, DevParticle *, T *const[3], const T *, const PCache<T> &, \
                                         double, int, double, int, uint64_t, uint32_t, const int *, double, double, \
                                         T *const[3], hipStream_t);                                                \
    template hipError_t slip_scatter<T>(const PParams &, DevParticle *, T *const[3], T *const[3], hipStream_t,      \
                                        const int *);                                                          \
    template hipError_t f_slip<T>(const PParams &, DevParticle *, T *const[3], double, hipStream_t, const int *);
KH_INST(double)
KH_INST(float)
#undef KH_INST

}  // namespace kh
