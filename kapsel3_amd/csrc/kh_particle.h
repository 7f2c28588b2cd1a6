// kh_particle.h -- device particle state and the per-point geometry of the
// smoothed profile (host + device).
#pragma once

#include "kh_common.h"

namespace kh {

// doubles per particle in one Output_udf frame (KH_FRAME_WORDS of the C ABI)
constexpr int FRAME_WORDS = 34;

// janus_propulsion values (enum JP, input.h; KH_JP_* of the C ABI)
constexpr int JP_MOTOR = 0, JP_SLIP = 1, JP_OBSTACLE = 2;

// Device copy of the members of Particle (variable.h:27-77) the step uses.
struct DevParticle {
    double x[3], x_prev[3], x_nopbc[3];
    double v[3], v_old[3], om[3], om_old[3];
    double q[4], q_old[4];  // (s, v0, v1, v2)
    double fr[3], fr_prev[3], tr[3], tr_prev[3];
    double fh[3], fh_prev[3], th[3], th_prev[3];
    double fs[3], fs_prev[3], ts[3], ts_prev[3];
    double v_slip[3], om_slip[3];
    double mass, mass_center[3], inertia[9];
    double smass, smc[3], sinertia[9], sdv[3], sdw[3];
    double mdfr[3];  // momentum_depend_fr: the step's LJ + gravity impulse (sheared runs)
    int    spec, pad_;
};

struct DevSpecies {
    double mass_ratio, imass, imoi;
    double janus_force[3], janus_torque[3];
    double slip_vel, slip_mode_half, c1, c2;
    int    propulsion, janus_axis;
};

// Everything the particle kernels need, passed by value.
struct PParams {
    int               nx, ny, nz;
    double            dx, L[3];
    double            radius, hxi, xi;
    int               rotation;
    int               np;
    int               np_domain;
    const int        *sek;  // packed Sekibun offsets
    const DevSpecies *species;
    int               lj_on, lj_powers, lj_cell_compat;
    double            lj_dia, lj_cutoff, epsilon;
    double            g;
    int               g_dir;
    double            gravity_on_fluid;
    int               motor;  // SW_JANUS_MOTOR
    double           *xs;     // np x 4: particle position, 1.0 for obstacles (compact copy for pair loops)
    int               direct_dist;  // every Sekibun offset is within half a box: |r| is the minimum-image distance
    // Lees-Edwards (the sheared step): the row of boxes above is displaced by
    // obl_g*LX... in x (obl_g*NY grid points) and moves obl_jump faster in x
    int               obl;
    double            obl_g, obl_jump;
    double            rho;     // RHO (the sheared step's stress sums)
    double           *pstress; // sheared runs: np x 6 per-particle stress partials
    // rigid composite bodies: body of each bead (nullptr: spherical
    // particles), and per bead the LJ force from partners of higher index
    const int        *bid;
    double           *flo;
    // the fused plain step: the phi scatter marks the 4-point vectors it
    // reaches (Occ, w == nullptr: not kept)
    Occ               occ;
};

// One rigid body (rigid.h's per-body state, input.cxx:1655-1695).
struct DevBody {
    double xG[3], xGn[3], xGp[3];            // xGs, xGs_nopbc, xGs_previous
    double vG[3], vGo[3], wG[3], wGo[3];     // velocityGs(_old), omegaGs(_old)
    double fGp[3], tGp[3], fGrp[3], tGrp[3]; // the previous step's forceGs, torqueGs, forceGrs, torqueGrs
    double M, Ib[3];                         // Rigid_Masses, principal moments (Rigid_Moments_body diagonal)
    int    first, last, free_, spec;         // beads [first, last), Rigid_motion, RigidID_Components
};
struct RigidArgs {
    DevBody      *B;
    int           nb;
    double       *GR;   // np x 3 GRvecs (space frame, from the centre to the bead)
    const double *GRb;  // np x 3 GRvecs_body
    int           compat;  // the reference's LJ torque bookkeeping (md_force.cxx:148-153)
};

// Host scalars of one sheared step's stress record (k_obl_record).
struct ShearRec {
    double time, rate, deg, strain, dt, eta;
    double ivolume;   // 1/(LX LY LZ)
    double ivol_dx3;  // DX^3/(LX LY LZ)
};

// Per-particle cache of the Sekibun points where phi is nonzero, filled by
// the phi scatter: entries [n*np_domain, n*np_domain + cnt[n]) hold, for
// particle n in Sekibun order, phi (field precision T), the grid index and the
// packed offset. The force gather and the u_p scatter then visit only these.
template <typename T>
struct PCache {
    T   *phi;
    int *idx;
    int *code;
    int *cnt;  // np entries
};

// Add_random_force_thermostat arguments (fluct.cxx:10-100)
struct Noise {
    int        on;
    uint64_t   seed;
    uint32_t   stream;
    const int *d_ts;
    double     niv, nio;
};

// One run's arguments of the particle launches of a batch step (kh_batch):
// the run's own values of what the per-run launches take; nz.d_ts is the
// run's step counter.
template <typename T>
struct SlotPart {
    PParams      pp;
    DevParticle *P;
    T           *phisum;
    PCache<T>    pc;
    T           *u[3], *up[3];
    Noise        nz;
    double      *traj;
    int          traj_base, traj_steps;
    double      *frames;
    int          frame_ts0, frame_every, frame_n;
};

__host__ __device__ inline int sek_pack(int i, int j, int k) { return (i + 128) | ((j + 128) << 8) | ((k + 128) << 16); }
__device__ __forceinline__ void sek_unpack(int c, int *o) {
    o[0] = (c & 255) - 128;
    o[1] = ((c >> 8) & 255) - 128;
    o[2] = ((c >> 16) & 255) - 128;
}

// Nint (macro.h:44-52)
__device__ __forceinline__ int nint(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }

// Distance0 (interaction.h:113-123): x12 = PBC(x2 - x1)
__device__ __forceinline__ double distance0(const PParams &pp, const double *x1, const double *x2, double *x12) {
    double dmy = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        x12[d] = x2[d] - x1[d];
        x12[d] -= (double)nint(x12[d] / pp.L[d]) * pp.L[d];
        dmy += x12[d] * x12[d];
    }
    return sqrt(dmy);
}

// Distance0_OBL (interaction.h:50-72): the y minimum image first; crossing
// it shifts x1 by the strain of the row of boxes it reaches
__device__ __forceinline__ double distance0_obl(const PParams &pp, const double *x1, const double *x2, double *x12) {
    double       dmy   = 0.0;
    const double signY = x2[1] - x1[1];
    x12[1]             = x2[1] - x1[1];
    x12[1] -= (double)nint(x12[1] / pp.L[1]) * pp.L[1];
    int sign = (int)(signY - x12[1]);
    if (!(sign == 0)) sign = sign / abs(sign);
    dmy += x12[1] * x12[1];
    x12[0] = x2[0] - (x1[0] + (double)sign * pp.obl_g * pp.L[1]);
    x12[0] -= (double)nint(x12[0] / pp.L[0]) * pp.L[0];
    dmy += x12[0] * x12[0];
    x12[2] = x2[2] - x1[2];
    x12[2] -= (double)nint(x12[2] / pp.L[2]) * pp.L[2];
    dmy += x12[2] * x12[2];
    return sqrt(dmy);
}

// H, Phi (profile.h:30-50)
__device__ __forceinline__ double prof_H(double x, double dx) { return x > 0 ? exp(-(dx / x) * (dx / x)) : 0.0; }
__device__ __forceinline__ double prof_phi(const PParams &pp, double x) {
    // inside the core H(x - a + xi/2) = 0, so phi = H/H = 1; outside the
    // interface H(a + xi/2 - x) = 0, so phi = 0: the same values the full
    // expression gives, without its two exp
    const double a1 = pp.radius + pp.hxi - x, a2 = x - pp.radius + pp.hxi;
    if (a2 <= 0.0) return 1.0;
    if (a1 <= 0.0) return 0.0;
    const double dmy = prof_H(a1, pp.dx);
    return dmy / (dmy + prof_H(a2, pp.dx));
}

// Phi in the field precision T: FP64 runs evaluate it in FP64; mixed runs in
// FP32 (the value is stored, scattered and weighted in FP32 anyway).
template <typename T>
__device__ __forceinline__ T prof_phi_t(const PParams &pp, T x) {
    if constexpr (sizeof(T) == 8) {
        return prof_phi(pp, x);
    } else {
        const float a1 = (float)(pp.radius + pp.hxi) - x, a2 = x - (float)(pp.radius - pp.hxi);
        if (a2 <= 0.0f) return 1.0f;
        if (a1 <= 0.0f) return 0.0f;
        const float dx = (float)pp.dx, t1 = dx / a1, t2 = dx / a2;
        const float h1 = __expf(-t1 * t1), h2 = __expf(-t2 * t2);
        return h1 / (h1 + h2);
    }
}

// Per-particle frame: Particle_cell (make_phi.h:114-139).
struct PCell {
    int    x_int[3];
    double residue[3];
};
__device__ __forceinline__ PCell particle_cell(const PParams &pp, const double *xp) {
    PCell        c;
    const double idx = 1. / pp.dx;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double dmy = xp[d] * idx;
        c.x_int[d]       = (int)dmy;
        c.residue[d]     = (dmy - (double)c.x_int[d]) * pp.dx;
    }
    return c;
}

// One Sekibun point: Relative_coord (make_phi.h:154-170) plus the distance
// Distance(r_mesh*DX, xp) that every particle loop of the reference uses.
struct PPoint {
    long long idx;  // real-space linear index (x, y, z)
    double    r[3]; // cell*dx - residue
    double    dist; // minimum-image |r_mesh*DX - xp|
    int       sign; // sheared runs: +1 / -1 when the point lies across the top / bottom boundary
};

// Relative_coord_check_stepover_Y (make_phi.h:229-268): on a sheared run a
// Sekibun point across the y boundary lies in the row of boxes displaced by
// sign*obl_g*NY grid points; returns that sign and sets r_mesh[0], r[0].
__device__ __forceinline__ int obl_stepover(const PParams &pp, const PCell &c, const int *cell, int *rm0, double *r0) {
    const int m1   = c.x_int[1] + cell[1];
    const int sign = m1 >= pp.ny ? 1 : (m1 < 0 ? -1 : 0);
    const double xo =
        fmod(c.x_int[0] - sign * pp.obl_g * pp.ny + c.residue[0] / pp.dx + cell[0] + 4 * pp.nx, (double)pp.nx);
    const int rm = (int)xo;
    *rm0         = min(max(rm, 0), pp.nx - 1);
    *r0          = (double)cell[0] * pp.dx - (xo - rm) * pp.dx;
    return sign;
}
// DIST = false leaves `dist` unset (the caller measures |r| itself); OBL:
// the sheared step's stepover geometry (compiled only into its kernels)
template <bool DIST = true, bool OBL = false>
__device__ __forceinline__ PPoint sek_point(const PParams &pp, const PCell &c, const double *xp, int code) {
    int cell[3];
    sek_unpack(code, cell);
    const int N[3] = {pp.nx, pp.ny, pp.nz};
    int       rm[3];
    double    x[3], x12[3];
    PPoint    q;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        // (x_int + cell + N) % N: x_int is in [0, N) and |cell| < N, so one
        // wrap suffices; the clamp keeps the index inside the grid even for
        // a non-finite particle state
        int m = c.x_int[d] + cell[d];
        m     = m < 0 ? m + N[d] : (m >= N[d] ? m - N[d] : m);
        rm[d] = min(max(m, 0), N[d] - 1);
        q.r[d] = (double)cell[d] * pp.dx - c.residue[d];
        x[d]   = rm[d] * pp.dx;
    }
    // Distance(r_mesh*DX, xp): when no offset reaches half a box the minimum
    // image is the offset itself, |cell*dx - residue| (equal up to rounding)
    if constexpr (DIST) {
        if (pp.direct_dist)
            q.dist = sqrt(q.r[0] * q.r[0] + q.r[1] * q.r[1] + q.r[2] * q.r[2]);
        else
            q.dist = distance0(pp, x, xp, x12);
    }
    q.sign = 0;
    if constexpr (OBL) {
        q.sign = obl_stepover(pp, c, cell, &rm[0], &q.r[0]);
        if constexpr (DIST) {
            x[0]   = rm[0] * pp.dx;
            q.dist = pp.direct_dist ? sqrt(q.r[0] * q.r[0] + q.r[1] * q.r[1] + q.r[2] * q.r[2])
                                    : distance0_obl(pp, x, xp, x12);
        }
    }
    q.idx  = ((long long)rm[0] * pp.ny + rm[1]) * pp.nz + rm[2];
    return q;
}

}  // namespace kh
