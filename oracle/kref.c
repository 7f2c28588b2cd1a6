/*
 * kref.c -- the oracle's SPM step (TEST INFRASTRUCTURE ONLY, see kref.h).
 *
 * A CPU FP64 restatement of Time_evolution_hydro (sp_3d_ns.cxx:41-194) for
 * SW_EQ == Navier_Stokes with spherical particles, on the reference's padded
 * field layout. Every function cites the reference lines it restates. Where
 * the reference's behaviour is replaced on purpose the comment says so.
 * kr_set_shear switches the run to the Lees-Edwards step
 * Time_evolution_hydro_OBL (sp_3d_ns.cxx:328-433). Replacements:
 *   - thermal noise: Philox4x32-10 keyed by (seed, stream, particle, step)
 *     instead of the process-global MT19937 (fluct.h:48-60), whose stream is
 *     not reproducible under OpenMP; the amplitude formula is the reference's.
 *   - LJ pairs: lj_cell_compat=1 reproduces the multiplicity with which the
 *     16-unit link cell (md_force.cxx:14, 56-101) counts each pair; 0 counts
 *     every pair once.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "kref_internal.h"

#define SQ(x) ((x) * (x))
#define KR_PI 3.14159265358979323846
static const double kT_snap_v = 1.439; /* input.cxx:309 */
static const double kT_snap_o = 1.22;  /* input.cxx:311 */
#define EPSILON_MP DBL_EPSILON         /* macro.h:36-38 */
#define TOL_MP (10.0 * EPSILON_MP)

/* one rigid body: rigid.h's per-body globals (input.cxx:1655-1695) */
typedef struct {
    double xG[3], xGn[3], xGp[3];    /* xGs, xGs_nopbc, xGs_previous          */
    double vG[3], vGo[3], wG[3], wGo[3];
    double fG[3], tG[3], fGr[3], tGr[3]; /* forceGs, torqueGs (hydro), forceGrs, torqueGrs (LJ, gravity) */
    double fGp[3], tGp[3], fGrp[3], tGrp[3];
    double M, Ib[3], Mom[9];         /* Rigid_Masses, Rigid_Moments_body diagonal, Rigid_Moments */
    int    free_;                    /* Rigid_motion "free" (1) or "fix" (0) */
    int    spec;                     /* RigidID_Components */
} kr_body;

struct kr_run {
    kh_params   P;
    kh_species *spec;
    int         nspec;
    int         NX, NY, NZ, HNX, HNY, HNZ_, NZ_, TRN_X, TRN_Y, TRN_Z;
    double      DX, L[3], WAVE[3];
    double      RADIUS, XI, HXI, SIGMA, RHO, IRHO, ETA, NU, DX3, kBT;
    double     *MASS, *IMASS, *MOI, *IMOI, *slip_mode_half;
    double      LJ_dia, A_R_cutoff;
    double      Gravity_on_fluid;
    int         sw_janus_motor, sw_janus_slip, fix_cell_any;
    size_t      M;
    double     *zeta[2], uk_dc[3], *u[3], *up[3], *f[3], *phi, *phi_sum, *f_ns0[2];
    int        *KX, *KY, *KZ;
    double     *K2, *IK2;
    int         np;
    Particle   *p;
    int         NP_domain;
    int        *sek;
    CTime       jikan;
    kr_fft3     fft;
    /* Lees-Edwards (SW_EQ == Shear_Navier_Stokes_Lees_Edwards) */
    int     le, le_ac;
    double  shear_rate, shear_freq, degree_oblique, shear_rate_eff, XYaspect, Ivolume;
    double *ucp[3], *w1, *w3[3];
    double *le_hf; /* per particle: Hydro_force, Hydro_force_new sums, volume, Itrace, force_x */
    double  le_stress_old, le_stress_new, le_reynolds, le_dev_lj, le_strain;
    double  le_record[KR_LE_RECORD];
    /* rigid composite bodies (SW_PT == rigid, kr_set_rigid) */
    int      nb;    /* bodies; 0: spherical particles */
    int     *bcum;  /* Rigid_Particle_Cumul: beads of body b are bcum[b] .. bcum[b+1]-1 */
    int     *bid;   /* Particle_RigidID */
    kr_body *body;
    double (*GR)[3], (*GRb)[3]; /* GRvecs, GRvecs_body */
};

#define IM(r, i, j, k) ((((size_t)(i) * (r)->NY + (j)) * (r)->NZ_) + (k))

/* Planted mistake (kr_set_mutation, kref.h): 0 in every use but the tests
 * that show a pin catches one. */
static int g_mut = KR_MUT_NONE;
void       kr_set_mutation(int m) { g_mut = (m > KR_MUT_NONE && m < KR_MUT_COUNT) ? m : KR_MUT_NONE; }

static void rigid_position(kr_run *r, int ab2);
static void rigid_velocity(kr_run *r, int ab2);

/* ------------------------------------------------------------------ */
/* small helpers: macro.h:44-52, interaction.h:113-132, lad3.h          */
/* ------------------------------------------------------------------ */
static int Nint(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }

static void distance0(const kr_run *r, const double *x1, const double *x2, double *r12, double *x12) {
    double dmy = 0.0;
    for (int d = 0; d < 3; d++) {
        x12[d] = x2[d] - x1[d];
        x12[d] -= (double)Nint(x12[d] / r->L[d]) * r->L[d];
        dmy += SQ(x12[d]);
    }
    *r12 = sqrt(dmy);
}

static double distance(const kr_run *r, const double *x1, const double *x2) {
    double r12, x12[3];
    distance0(r, x1, x2, &r12, x12);
    return r12;
}

static int less_than_mp(double a, double b) { /* macro.h:164-167 */
    double eps = (fmax(fabs(a), fabs(b)) + 10.0) * EPSILON_MP;
    return (b - a) > eps;
}

static void v_cross(double *c, const double *a, const double *b, double alpha) { /* lad3.h:390-395 */
    c[0] = alpha * (a[1] * b[2] - a[2] * b[1]);
    c[1] = alpha * (a[2] * b[0] - a[0] * b[2]);
    c[2] = alpha * (a[0] * b[1] - a[1] * b[0]);
}

/* quaternion.h:160-168, 178-185 (qtn_prod, qtn_qvq_prod) */
static void qtn_prod(quaternion *q, const quaternion *qa, const quaternion *qb, double alpha) {
    q->s    = alpha * (qa->s * qb->s - (qa->v[0] * qb->v[0] + qa->v[1] * qb->v[1] + qa->v[2] * qb->v[2]));
    q->v[0] = alpha * (qa->s * qb->v[0] + qb->s * qa->v[0] + qa->v[1] * qb->v[2] - qa->v[2] * qb->v[1]);
    q->v[1] = alpha * (qa->s * qb->v[1] + qb->s * qa->v[1] + qa->v[2] * qb->v[0] - qa->v[0] * qb->v[2]);
    q->v[2] = alpha * (qa->s * qb->v[2] + qb->s * qa->v[2] + qa->v[0] * qb->v[1] - qa->v[1] * qb->v[0]);
}
static void qtn_qvq_prod(double *qvq, const quaternion *q, const double *v) {
    double qs2   = q->s * q->s - (q->v[0] * q->v[0] + q->v[1] * q->v[1] + q->v[2] * q->v[2]);
    double qdotp = q->v[0] * v[0] + q->v[1] * v[1] + q->v[2] * v[2];
    qvq[0]       = (qs2 * v[0] + 2.0 * (qdotp * q->v[0] + q->s * (q->v[1] * v[2] - q->v[2] * v[1])));
    qvq[1]       = (qs2 * v[1] + 2.0 * (qdotp * q->v[1] + q->s * (q->v[2] * v[0] - q->v[0] * v[2])));
    qvq[2]       = (qs2 * v[2] + 2.0 * (qdotp * q->v[2] + q->s * (q->v[0] * v[1] - q->v[1] * v[0])));
}
/* rigid_body.cxx:10-24 */
static void body2space(double *out, const double *in, const quaternion *q) { qtn_qvq_prod(out, q, in); }
static void space2body(double *out, const double *in, const quaternion *q) {
    quaternion c = *q;
    c.v[0]       = -c.v[0];
    c.v[1]       = -c.v[1];
    c.v[2]       = -c.v[2];
    qtn_qvq_prod(out, &c, in);
}
/* rigid_body.cxx:157-168, SPACE_FRAME: dq/dt = 1/2 (0,w) o q */
static void qdot_space(quaternion *dqdt, const quaternion *q, const double *omega) {
    quaternion qw;
    qw.s    = 0.0;
    qw.v[0] = omega[0];
    qw.v[1] = omega[1];
    qw.v[2] = omega[2];
    qtn_prod(dqdt, &qw, q, 0.5);
}
static void qtn_add(quaternion *q, const quaternion *qb, double alpha) { /* quaternion.h:131-136 */
    q->s += qb->s * alpha;
    q->v[0] += qb->v[0] * alpha;
    q->v[1] += qb->v[1] * alpha;
    q->v[2] += qb->v[2] * alpha;
}
static void qtn_normalize(quaternion *q) { /* quaternion.h:260-268 */
    double qnorm = sqrt(q->s * q->s + q->v[0] * q->v[0] + q->v[1] * q->v[1] + q->v[2] * q->v[2]);
    if (!((qnorm > TOL_MP) || (-TOL_MP > qnorm))) {
        fprintf(stderr, "kref: zero-norm quaternion\n");
        abort();
    }
    qnorm = 1.0 / qnorm;
    q->s *= qnorm;
    q->v[0] *= qnorm;
    q->v[1] *= qnorm;
    q->v[2] *= qnorm;
}

/* ------------------------------------------------------------------ */
/* profile: profile.h:30-50, 175-184; profile.cxx:22-65                 */
/* ------------------------------------------------------------------ */
static double H(double x, double dx) { return x > 0 ? exp(-SQ(dx / x)) : 0; }

double kr_phi(double x, double radius, double dx, double xi) {
    double hxi = 0.5 * xi;
    double dmy = H(radius + hxi - x, dx);
    return dmy / (dmy + H(x - radius + hxi, dx));
}

static double Phi(const kr_run *r, double x) { return kr_phi(x, r->RADIUS, r->DX, r->XI); }

static double DPhi_compact_sin_norm(const kr_run *r, double x) {
    const double hxi = r->XI * 0.5, ixi = 1.0 / r->XI;
    double       dmy_x = r->RADIUS - x;
    return (fabs(dmy_x) < hxi) ? cos(KR_PI * dmy_x * ixi) : 0.0;
}

static void particle_domain(kr_run *r) {
    const kh_params *P  = &r->P;
    int              nh = (int)P->a + (int)P->a_xi + 2;
    double           MAX_RADIUS = r->RADIUS + r->HXI + 1.7321 * r->DX;
    int              cap        = (2 * nh + 1) * (2 * nh + 1) * (2 * nh + 1);
    r->sek                      = (int *)malloc(sizeof(int) * 3 * (size_t)cap);
    int mesh                    = 0;
    for (int i = -nh; i <= nh; i++)
        for (int j = -nh; j <= nh; j++)
            for (int k = -nh; k <= nh; k++) {
                double dmy = sqrt(SQ((double)i * r->DX) + SQ((double)j * r->DX) + SQ((double)k * r->DX));
                if (dmy < MAX_RADIUS) {
                    r->sek[3 * mesh + 0] = i;
                    r->sek[3 * mesh + 1] = j;
                    r->sek[3 * mesh + 2] = k;
                    mesh++;
                }
            }
    r->NP_domain = mesh;
}

int kr_sekibun_count(const kr_run *r) { return r->NP_domain; }
int kr_sekibun_cell(const kr_run *r, int *out3) {
    memcpy(out3, r->sek, sizeof(int) * 3 * (size_t)r->NP_domain);
    return r->NP_domain;
}

/* make_phi.h:114-139 */
static void particle_cell(const kr_run *r, const double *xp, int *x_int, double *residue) {
    const double idx = 1. / r->DX;
    for (int d = 0; d < 3; d++) {
        double dmy = xp[d] * idx;
        x_int[d]   = (int)dmy;
        residue[d] = (dmy - (double)x_int[d]) * r->DX;
    }
}
/* make_phi.h:154-170 */
static void relative_coord(const kr_run *r, const int *cell, const int *x_int, const double *residue, int *r_mesh,
                           double *rv) {
    const int N[3] = {r->NX, r->NY, r->NZ};
    for (int d = 0; d < 3; d++) {
        r_mesh[d] = (x_int[d] + cell[d] + N[d]) % N[d];
        rv[d]     = (double)cell[d] * r->DX - residue[d];
    }
}
/* make_phi.h:388-407 */
static void angular2v(const kr_run *r, const double *omega, const double *rv, double *v) {
    if (r->P.rotation) {
        v[0] = omega[1] * rv[2] - omega[2] * rv[1];
        v[1] = omega[2] * rv[0] - omega[0] * rv[2];
        v[2] = omega[0] * rv[1] - omega[1] * rv[0];
        if (g_mut == KR_MUT_OMEGA_R_SIGN)
            for (int d = 0; d < 3; d++) v[d] = -v[d];
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* spectral operators                                                 */
/* ------------------------------------------------------------------ */
/* fft_wrapper.cxx:45-76 with Calc_K*_Ooura (fft_wrapper.h:719-729) */
static void init_K(kr_run *r) {
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                int    kx = (i > r->HNX) ? i - r->NX : i;
                int    ky = (j > r->HNY) ? j - r->NY : j;
                int    kz = k / 2;
                r->KX[im] = kx;
                r->KY[im] = ky;
                r->KZ[im] = kz;
                r->K2[im] = SQ(r->WAVE[0] * (double)kx) + SQ(r->WAVE[1] * (double)ky) + SQ(r->WAVE[2] * (double)kz);
                r->IK2[im] = (r->K2[im] > 0.0) ? 1. / r->K2[im] : 0.0;
            }
}

static void A2a_k(kr_run *r, double *a) { kr_fft3_forward(&r->fft, a); }
static void A_k2a(kr_run *r, double *a) { kr_fft3_inverse(&r->fft, a); }
void        kr_A2a_k(kr_run *r, double *a) { A2a_k(r, a); }
void        kr_A_k2a(kr_run *r, double *a) { A_k2a(r, a); }

/* Truncate_two_third_rule_ooura (fft_wrapper.h:754-796): the union of its
 * four Truncate_general boxes is {kz >= TRN_Z} u {|ky| >= TRN_Y} u {|kx| >= TRN_X}. */
static int truncated(const kr_run *r, int i, int j, int k) {
    if (k >= 2 * r->TRN_Z) return 1;
    if (j >= r->TRN_Y && j <= r->NY - r->TRN_Y) return 1;
    if (i >= r->TRN_X && i <= r->NX - r->TRN_X) return 1;
    return 0;
}
static void truncate_two_third(kr_run *r, double *a) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++)
                if (truncated(r, i, j, k)) a[IM(r, i, j, k)] = 0.0;
}

/* Zeta_k2u_k (fft_wrapper.cxx:591-688) */
static void zeta_k2u_k(kr_run *r, double **zeta, const double *uk_dc, double **u) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int kk = 0; kk < r->HNZ_; kk++) {
                size_t im = IM(r, i, j, 2 * kk);
                double kx = r->WAVE[0] * r->KX[im], ky = r->WAVE[1] * r->KY[im], kz = r->WAVE[2] * r->KZ[im];
                double ik2 = r->IK2[im];
                double d1r = zeta[0][im], d1i = zeta[0][im + 1], d2r = zeta[1][im], d2i = zeta[1][im + 1];
                double ore[3], oim[3];
                if (r->KX[im] != 0) {
                    ore[0] = -(1. / kx) * (ky * d1r + kz * d2r);
                    oim[0] = -(1. / kx) * (ky * d1i + kz * d2i);
                    ore[1] = d1r;
                    oim[1] = d1i;
                    ore[2] = d2r;
                    oim[2] = d2i;
                } else if (r->KY[im] != 0) {
                    double dmy = -(kz / ky);
                    ore[0]     = d2r;
                    oim[0]     = d2i;
                    ore[1]     = dmy * d1r;
                    oim[1]     = dmy * d1i;
                    ore[2]     = d1r;
                    oim[2]     = d1i;
                } else {
                    ore[0] = d1r;
                    oim[0] = d1i;
                    ore[1] = d2r;
                    oim[1] = d2i;
                    ore[2] = 0.0;
                    oim[2] = 0.0;
                }
                kx *= ik2;
                ky *= ik2;
                kz *= ik2;
                u[0][im]     = -(ky * oim[2] - kz * oim[1]);
                u[0][im + 1] = (ky * ore[2] - kz * ore[1]);
                u[1][im]     = -(kz * oim[0] - kx * oim[2]);
                u[1][im + 1] = (kz * ore[0] - kx * ore[2]);
                u[2][im]     = -(kx * oim[1] - ky * oim[0]);
                u[2][im + 1] = (kx * ore[1] - ky * ore[0]);
            }
    u[0][0] = uk_dc[0];
    u[1][0] = uk_dc[1];
    u[2][0] = uk_dc[2];
}

/* U_k2zeta_k (fft_wrapper.cxx:407-476) */
static void u_k2zeta_k(kr_run *r, double **u, double **zeta, double *uk_dc) {
    uk_dc[0] = u[0][0];
    uk_dc[1] = u[1][0];
    uk_dc[2] = u[2][0];
    if (g_mut == KR_MUT_UKDC_MEAN)
        for (int d = 0; d < 3; d++) uk_dc[d] /= (double)r->NX * r->NY * r->NZ;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int kk = 0; kk < r->HNZ_; kk++) {
                size_t im = IM(r, i, j, 2 * kk);
                double ks[3] = {r->KX[im] * r->WAVE[0], r->KY[im] * r->WAVE[1], r->KZ[im] * r->WAVE[2]};
                double ore[3], oim[3];
                ore[0] = -(ks[1] * u[2][im + 1] - ks[2] * u[1][im + 1]);
                oim[0] = (ks[1] * u[2][im] - ks[2] * u[1][im]);
                ore[1] = -(ks[2] * u[0][im + 1] - ks[0] * u[2][im + 1]);
                oim[1] = (ks[2] * u[0][im] - ks[0] * u[2][im]);
                ore[2] = -(ks[0] * u[1][im + 1] - ks[1] * u[0][im + 1]);
                oim[2] = (ks[0] * u[1][im] - ks[1] * u[0][im]);
                int a0, a1;
                if (r->KX[im] != 0) {
                    a0 = 1;
                    a1 = 2;
                } else if (r->KY[im] != 0) {
                    a0 = 2;
                    a1 = 0;
                } else {
                    a0 = 0;
                    a1 = 1;
                }
                zeta[0][im]     = ore[a0];
                zeta[0][im + 1] = oim[a0];
                zeta[1][im]     = ore[a1];
                zeta[1][im + 1] = oim[a1];
            }
}

/* U2advection_k (operate_omega.cxx:12-85) with Omega_k2zeta_k (fft_wrapper.cxx:227-263) */
static void u2advection_k(kr_run *r, double **u, double **adv) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                double u1 = u[0][im], u2 = u[1][im], u3 = u[2][im];
                u[0][im]   = g_mut == KR_MUT_ADV_SIGN ? -(u2 * u3) : u2 * u3;
                u[1][im]   = u3 * u1;
                u[2][im]   = u1 * u2;
                adv[0][im] = (u2 - u3) * (u2 + u3);
                adv[1][im] = (u3 - u1) * (u3 + u1);
            }
    for (int d = 0; d < 3; d++) A2a_k(r, u[d]);
    A2a_k(r, adv[0]);
    A2a_k(r, adv[1]);
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im      = IM(r, i, j, k);
                double k1      = r->KX[im] * r->WAVE[0];
                double k2      = r->KY[im] * r->WAVE[1];
                double k3      = r->KZ[im] * r->WAVE[2];
                double u2u3    = u[0][im];
                double u3u1    = u[1][im];
                double u1u2    = u[2][im];
                double u22_u32 = adv[0][im];
                double u32_u12 = adv[1][im];
                double u12_u22 = -u32_u12 - u22_u32;
                double k1k2 = k1 * k2, k2k3 = k2 * k3, k3k1 = k3 * k1;
                u[0][im] = -(k2k3 * u22_u32 + (k3 - k2) * (k3 + k2) * u2u3 + k3k1 * u1u2 - k1k2 * u3u1);
                u[1][im] = -(k3k1 * u32_u12 + (k1 - k3) * (k1 + k3) * u3u1 + k1k2 * u2u3 - k2k3 * u1u2);
                u[2][im] = -(k1k2 * u12_u22 + (k2 - k1) * (k2 + k1) * u1u2 + k2k3 * u3u1 - k3k1 * u2u3);
            }
    /* Omega_k2zeta_k */
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                if (r->KX[im] != 0) {
                    adv[0][im] = u[1][im];
                    adv[1][im] = u[2][im];
                } else if (r->KY[im] != 0) {
                    adv[0][im] = u[2][im];
                    adv[1][im] = u[0][im];
                } else {
                    adv[0][im] = u[0][im];
                    adv[1][im] = u[1][im];
                }
            }
}

/* Zeta_k2advection_k (operate_omega.cxx:87-98) */
static void zeta_k2advection_k(kr_run *r, double **zeta, double *uk_dc, double **adv) {
    truncate_two_third(r, zeta[0]);
    truncate_two_third(r, zeta[1]);
    zeta_k2u_k(r, zeta, uk_dc, r->u);
    for (int d = 0; d < 3; d++) A_k2a(r, r->u[d]);
    u2advection_k(r, r->u, adv);
}

/* NS_solver_slavedEuler (fluid_solver.cxx:348-374); the kept-mode boxes
 * ijk_range_two_third_filter (fft_wrapper.cxx:172-183). */
static void slaved_euler_update(kr_run *r) {
    const double dmy0 = -r->NU * r->jikan.dt_fluid;
    const double dmy1 = 1. / r->NU;
    const int    TX = r->TRN_X, TY = r->TRN_Y, TZ = r->TRN_Z;
    int          ib[4][2] = {{0, TX - 1}, {0, TX - 1}, {r->NX - TX + 1, r->NX - 1}, {r->NX - TX + 1, r->NX - 1}};
    int          jb[4][2] = {{0, TY - 1}, {r->NY - TY + 1, r->NY - 1}, {0, TY - 1}, {r->NY - TY + 1, r->NY - 1}};
    for (int n = 0; n < 4; n++) {
#pragma omp parallel for
        for (int i = ib[n][0]; i <= ib[n][1]; i++)
            for (int j = jb[n][0]; j <= jb[n][1]; j++)
                for (int k = 0; k <= 2 * TZ - 1; k++) {
                    size_t im  = IM(r, i, j, k);
                    double dmy = exp(dmy0 * r->K2[im]) - 1.;
                    r->zeta[0][im] += dmy * (r->zeta[0][im] - dmy1 * r->IK2[im] * r->f_ns0[0][im]);
                    r->zeta[1][im] += dmy * (r->zeta[1][im] - dmy1 * r->IK2[im] * r->f_ns0[1][im]);
                }
    }
}
static void ns_solver_slaved_euler(kr_run *r) {
    zeta_k2advection_k(r, r->zeta, r->uk_dc, r->f_ns0);
    slaved_euler_update(r);
}

/* Solenoidal_uk (operate_omega.cxx:171-195) */
static void solenoidal_uk(kr_run *r, double **u) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im  = IM(r, i, j, k);
                double kx  = r->KX[im] * r->WAVE[0];
                double ky  = r->KY[im] * r->WAVE[1];
                double kz  = r->KZ[im] * r->WAVE[2];
                double dmy = r->IK2[im] * (u[0][im] * kx + u[1][im] * ky + u[2][im] * kz);
                u[0][im] -= dmy * kx;
                u[1][im] -= dmy * ky;
                u[2][im] -= dmy * kz;
            }
}

/* Solenoidal_u (operate_omega.h:105-112) */
static void solenoidal_u(kr_run *r, double **u) {
    for (int d = 0; d < 3; d++) A2a_k(r, u[d]);
    for (int d = 0; d < 3; d++) truncate_two_third(r, u[d]);
    solenoidal_uk(r, u);
    for (int d = 0; d < 3; d++) A_k2a(r, u[d]);
}

/* Zeta_k2u (operate_omega.h:134-137) */
static void zeta_k2u(kr_run *r, double **zeta, double *uk_dc, double **u) {
    zeta_k2u_k(r, zeta, uk_dc, u);
    for (int d = 0; d < 3; d++) A_k2a(r, u[d]);
}

/* U2zeta_k (operate_omega.h:208-211) */
static void u2zeta_k(kr_run *r, double **zeta, double *uk_dc, double **u) {
    for (int d = 0; d < 3; d++) A2a_k(r, u[d]);
    u_k2zeta_k(r, u, zeta, uk_dc);
}

/* ------------------------------------------------------------------ */
/* particle <-> grid                                                   */
/* ------------------------------------------------------------------ */
static void reset_field(kr_run *r, double *a) { memset(a, 0, sizeof(double) * r->M); } /* Reset_phi/Reset_u */

/* Make_phi_particle_sum (make_phi.cxx:214-260) */
static void make_phi_particle_sum(kr_run *r, double *phi, double *phi_sum) {
    for (int n = 0; n < r->np; n++) { /* serial: fixed summation order */
        double xp[3] = {r->p[n].x[0], r->p[n].x[1], r->p[n].x[2]};
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            double dmy = distance(r, x, xp);
            phi_sum[IM(r, r_mesh[0], r_mesh[1], r_mesh[2])] += Phi(r, dmy);
        }
    }
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                phi[im]   = fmin(phi_sum[im], 1.0);
            }
}

/* Make_u_particle_sum (make_phi.cxx:313-354) */
static void make_u_particle_sum(kr_run *r, double **up, const double *phi_sum) {
    for (int n = 0; n < r->np; n++) {
        const Particle *pn = &r->p[n];
        double          xp[3], vp[3], om[3];
        for (int d = 0; d < 3; d++) {
            xp[d] = pn->x[d];
            vp[d] = pn->v[d];
            om[d] = pn->omega[d];
        }
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3], v_rot[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            size_t im      = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            double dmy     = distance(r, x, xp);
            double dmy_phi = Phi(r, dmy) / (g_mut == KR_MUT_UP_NO_MAX ? 1.0 : fmax(phi_sum[im], 1.0));
            angular2v(r, om, rv, v_rot);
            up[0][im] += ((vp[0] + v_rot[0]) * dmy_phi);
            up[1][im] += ((vp[1] + v_rot[1]) * dmy_phi);
            up[2][im] += ((vp[2] + v_rot[2]) * dmy_phi);
        }
    }
}

/* Calc_f_hydro_correct_precision (md_force.cxx:383-506), non-rigid branch */
static void calc_f_hydro(kr_run *r, const double *phi_sum, double **u) {
    const double dmy = (g_mut == KR_MUT_FH_SIGN ? r->DX3 * r->RHO : -r->DX3 * r->RHO) / r->jikan.dt_fluid;
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *pn = &r->p[n];
        double    xp[3], vp[3], om[3], force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) {
            xp[d] = pn->x[d];
            vp[d] = pn->v[d];
            om[d] = pn->omega[d];
        }
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3], v_rot[3], fp[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            double dmyR = distance(r, x, xp);
            angular2v(r, om, rv, v_rot);
            size_t im      = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            double dmy_phi = Phi(r, dmyR) / (g_mut == KR_MUT_FH_NO_MAX ? 1.0 : fmax(phi_sum[im], 1.0));
            for (int d = 0; d < 3; d++) {
                fp[d] = ((vp[d] + v_rot[d]) - u[d][im]) * dmy_phi;
                force[d] += fp[d];
            }
            torque[0] += (rv[1] * fp[2] - rv[2] * fp[1]);
            torque[1] += (rv[2] * fp[0] - rv[0] * fp[2]);
            torque[2] += (rv[0] * fp[1] - rv[1] * fp[0]);
        }
        for (int d = 0; d < 3; d++) {
            pn->f_hydro[d] = (dmy * force[d]);
            pn->f_slip[d]  = 0.0;
        }
        if (r->P.rotation)
            for (int d = 0; d < 3; d++) {
                pn->torque_hydro[d] = (dmy * torque[d]);
                pn->torque_slip[d]  = 0.0;
            }
    }
    /* rigid: forceGs += sum of the beads' f_hydro, torqueGs += sum of
     * (GR + r) x f over their points = GR x f_hydro + torque_hydro per bead
     * (md_force.cxx:467-473, 496-505), in bead order */
    for (int n = 0; n < r->np && r->nb; n++) {
        kr_body        *b = &r->body[r->bid[n]];
        const Particle *pn = &r->p[n];
        double          gxf[3];
        v_cross(gxf, r->GR[n], pn->f_hydro, 1.0);
        for (int d = 0; d < 3; d++) {
            b->fG[d] += pn->f_hydro[d];
            b->tG[d] += gxf[d] + pn->torque_hydro[d];
        }
    }
}

/* Make_f_particle_dt_nonsole (f_particle.h:35-54) */
static void make_f_particle_nonsole(kr_run *r, double **f, double **u, double **up, const double *phi) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                f[0][im]  = up[0][im] - phi[im] * u[0][im];
                f[1][im]  = up[1][im] - phi[im] * u[1][im];
                f[2][im]  = up[2][im] - phi[im] * u[2][im];
            }
}
/* Add_f_particle (f_particle.h:68-84) */
static void add_f_particle(kr_run *r, double **u, double **f) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                u[0][im] += f[0][im];
                u[1][im] += f[1][im];
                u[2][im] += f[2][im];
            }
}

/* ------------------------------------------------------------------ */
/* particle forces and integrators                                      */
/* ------------------------------------------------------------------ */
/* Lennard_Jones_f (interaction.h:160-199), powers 0..2 */
static double lennard_jones_f(const kr_run *r, double x) {
    double sigma = r->LJ_dia, eps = r->P.epsilon, dmy = sigma / x, answer = 0.0;
    if (r->P.lj_powers == 0) {
        dmy    = SQ(dmy) * SQ(dmy) * SQ(dmy);
        answer = (24. * eps) / SQ(x) * (2.0 * SQ(dmy) - dmy);
    } else if (r->P.lj_powers == 1) {
        dmy    = SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy);
        answer = (48. * eps) / SQ(x) * (2.0 * SQ(dmy) - dmy);
    } else if (r->P.lj_powers == 2) {
        dmy    = SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy) * SQ(dmy);
        answer = (72. * eps) / SQ(x) * (2.0 * SQ(dmy) - dmy);
    }
    return answer;
}

/* Multiplicity with which the 16-unit link cell of
 * Calc_f_Lennard_Jones_shear_cap_primitive_lnk (md_force.cxx:34-196) visits
 * the pair (i,j): per axis 3 if there is one cell, 2 if there are two and the
 * particles sit in different cells, 1 otherwise; 0 if the cells are not
 * neighbours. */
static int lj_multiplicity(const kr_run *r, const double *xi, const double *xj) {
    if (!r->P.lj_cell_compat) return 1;
    const int N[3] = {r->NX, r->NY, r->NZ};
    int       m    = 1;
    for (int d = 0; d < 3; d++) {
        int    lc = N[d] / 16;
        double cl = 16 * r->DX;
        int    ci = (int)(xi[d] / cl), cj = (int)(xj[d] / cl);
        int    md;
        if (lc == 1)
            md = 3;
        else if (lc == 2)
            md = (ci == cj) ? 1 : 2;
        else {
            int dd = ((ci - cj) % lc + lc) % lc;
            md     = (dd == 0 || dd == 1 || dd == lc - 1) ? 1 : 0;
        }
        m *= md;
    }
    return m;
}

static void calc_f_lennard_jones(kr_run *r) {
    const double pair_cutoff = r->A_R_cutoff * r->LJ_dia;
    /* every unordered pair, i > j, accumulated in a fixed order */
    for (int i = 0; i < r->np; i++)
        for (int j = 0; j < i; j++) {
            Particle *pi = &r->p[i], *pj = &r->p[j];
            if (r->spec[pi->spec].propulsion == KH_JP_OBSTACLE && r->spec[pj->spec].propulsion == KH_JP_OBSTACLE)
                continue; /* obstacle_chain, interaction.h:29-35 */
            if (r->nb && r->bid[i] == r->bid[j]) continue; /* rigid_chain, interaction.h:18-24 */
            double rij, rv[3];
            distance0(r, pi->x, pj->x, &rij, rv);
            if (rij < pair_cutoff) {
                int mult = lj_multiplicity(r, pi->x, pj->x);
                double dmy_r = fmin(DBL_MAX / rij, lennard_jones_f(r, rij));
                for (int c = 0; c < mult; c++) {
                    double fi[3];
                    for (int d = 0; d < 3; d++) {
                        fi[d] = dmy_r * (-rv[d]);
                        pi->fr[d] += fi[d];
                        pj->fr[d] -= fi[d];
                    }
                    if (r->nb) { /* forceGrs, torqueGrs (md_force.cxx:134-153) */
                        kr_body *bi = &r->body[r->bid[i]], *bj = &r->body[r->bid[j]];
                        double   ti[3], tj[3];
                        v_cross(ti, r->GR[i], fi, 1.0);
                        /* the reference adds GR_j x f_i to body j, whose bead takes -f_i;
                         * lj_cell_compat = 0 gives it GR_j x (-f_i) */
                        v_cross(tj, r->GR[j], fi, r->P.lj_cell_compat ? 1.0 : -1.0);
                        for (int d = 0; d < 3; d++) {
                            bi->fGr[d] += fi[d];
                            bj->fGr[d] -= fi[d];
                            bi->tGr[d] += ti[d];
                            bj->tGr[d] += tj[d];
                        }
                    }
                }
            }
        }
}

/* Add_f_gravity (md_force.cxx:306-324) */
static void add_f_gravity(kr_run *r) {
    if (r->nb) { /* rigid branch: on each body's volume */
        for (int k = 0; k < r->nb; k++) {
            kr_body     *b     = &r->body[k];
            const double ratio = r->spec[b->spec].mass_ratio;
            b->fGr[r->P.g_dir] -= r->P.g * r->RHO * (b->M / (ratio * r->RHO)) * (ratio - 1.0);
        }
        return;
    }
    for (int n = 0; n < r->np; n++)
        r->p[n].fr[r->P.g_dir] -= r->Gravity_on_fluid * (r->spec[r->p[n].spec].mass_ratio - 1.0);
}

/* Force (particle_solver.h:88-111) for the spherical, wall-free case */
static void force(kr_run *r) {
    if (r->P.lj_truncate >= 0) calc_f_lennard_jones(r);
    if (r->P.g != 0.0) add_f_gravity(r);
}

/* reset_Forces (particle_solver.cxx:11-37) */
static void reset_forces(kr_run *r) {
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        for (int d = 0; d < 3; d++) {
            p->fr_previous[d]           = p->fr[d];
            p->fr[d]                    = 0.0;
            p->f_hydro_previous[d]      = p->f_hydro[d];
            p->f_hydro[d]               = 0.0;
            p->f_slip_previous[d]       = p->f_slip[d];
            p->f_slip[d]                = 0.0;
            p->torque_r_previous[d]     = p->torque_r[d];
            p->torque_r[d]              = 0.0;
            p->torque_hydro_previous[d] = p->torque_hydro[d];
            p->torque_hydro[d]          = 0.0;
            p->torque_slip_previous[d]  = p->torque_slip[d];
            p->torque_slip[d]           = 0.0;
        }
    }
}

static int is_obstacle(const kr_run *r, const Particle *p) { return r->spec[p->spec].propulsion == KH_JP_OBSTACLE; }

/* MD_solver_orientation_Euler / _AB2 (particle_rotation_solver.h:23-29, 72-81) */
static void orientation_update(const kr_run *r, Particle *p, int ab2) {
    if (!ab2) {
        quaternion dqdt;
        p->q_old = p->q;
        qdot_space(&dqdt, &p->q, p->omega);
        qtn_add(&p->q, &dqdt, r->jikan.dt_md);
        qtn_normalize(&p->q);
    } else {
        quaternion dqdt, dqdt_old;
        qdot_space(&dqdt, &p->q, p->omega);
        qdot_space(&dqdt_old, &p->q_old, p->omega_old);
        p->q_old = p->q;
        qtn_add(&p->q, &dqdt, 3.0 * r->jikan.hdt_md);
        qtn_add(&p->q, &dqdt_old, -r->jikan.hdt_md);
        qtn_normalize(&p->q);
    }
}

/* MD_solver_position_Euler / _AB2 (particle_solver.cxx:39-85) */
static void position_update(kr_run *r, int ab2) {
    if (r->nb) { /* SW_PT == rigid */
        rigid_position(r, ab2);
        return;
    }
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        if (is_obstacle(r, p)) continue;
        for (int d = 0; d < 3; d++) {
            p->x_previous[d] = p->x[d];
            double delta_x   = ab2 ? r->jikan.hdt_md * (3.0 * p->v[d] - p->v_old[d]) : r->jikan.dt_md * p->v[d];
            p->x_nopbc[d] += delta_x;
            p->x[d] += delta_x;
        }
        for (int d = 0; d < 3; d++) p->x[d] = fmod(p->x[d] + r->L[d], r->L[d]); /* PBC, periodic_boundary.h:9-15 */
        if (r->P.rotation) orientation_update(r, p, ab2);
    }
}

/* self_propulsion (particle_solver.cxx:88-101) */
static void self_propulsion(const kr_run *r, const Particle *p, double *f, double *t) {
    f[0] = f[1] = f[2] = t[0] = t[1] = t[2] = 0.0;
    if (r->sw_janus_motor && r->spec[p->spec].propulsion == KH_JP_MOTOR) {
        body2space(f, r->spec[p->spec].janus_force, &p->q);
        body2space(t, r->spec[p->spec].janus_torque, &p->q);
    }
}

/* MD_solver_velocity_Euler (103-142) and MD_solver_velocity_AB2_hydro (239-281) */
static void velocity_update(kr_run *r, int ab2) {
    force(r);
    if (r->nb) { /* calc_Rigid_VOGs + set_Particle_Velocities (particle_solver.cxx:137-140, 274-277) */
        rigid_velocity(r, ab2);
        reset_forces(r);
        return;
    }
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p    = &r->p[n];
        int       s    = p->spec;
        double    h    = ab2 ? r->jikan.hdt_md : r->jikan.dt_md;
        double    dmy  = h * r->IMASS[s];
        double    dmyr = h * r->IMOI[s];
        if (!is_obstacle(r, p)) {
            double sf[3], st[3];
            self_propulsion(r, p, sf, st);
            for (int d = 0; d < 3; d++) {
                p->v_old[d]     = p->v[d];
                p->omega_old[d] = p->omega[d];
                if (!ab2) {
                    p->v[d] += (dmy * (p->f_hydro[d] + p->fr[d] + sf[d]));
                    p->omega[d] += (dmyr * (p->torque_hydro[d] + p->torque_r[d] + st[d]));
                } else {
                    const double w = g_mut == KR_MUT_AB2_FH_WEIGHT ? 1.0 : 2.0;
                    p->v[d] += dmy * (w * (p->f_hydro[d] + sf[d]) + p->fr[d] + p->fr_previous[d]);
                    p->omega[d] +=
                        dmyr * (w * (p->torque_hydro[d] + st[d]) + p->torque_r[d] + p->torque_r_previous[d]);
                }
            }
        } else {
            for (int d = 0; d < 3; d++) {
                p->v_old[d]     = p->v[d];
                p->v[d]         = 0.0;
                p->omega_old[d] = p->omega[d];
                p->omega[d]     = 0.0;
            }
        }
    }
    reset_forces(r);
}

/* ------------------------------------------------------------------ */
/* rigid composite bodies (SW_PT == rigid)                              */
/* ------------------------------------------------------------------ */
static void pbc3(const kr_run *r, double *x) { /* PBC (periodic_boundary.h:9-15) */
    for (int d = 0; d < 3; d++) x[d] = fmod(x[d] + r->L[d], r->L[d]);
}

/* MD_solver_omega_Euler / _AB2 (particle_rotation_solver.h:106-184):
 * Euler's equations in the body frame with the body's principal moments */
static void omega_euler(double *w, const double *tq, const double *Ib, const quaternion *q, int free_, double dt) {
    double wb[3], tb[3];
    space2body(wb, w, q);
    space2body(tb, tq, q);
    double nw[3] = {wb[0], wb[1], wb[2]};
    if (free_) {
        nw[0] += (dt / Ib[0]) * (tb[0] + (Ib[1] - Ib[2]) * wb[1] * wb[2]);
        nw[1] += (dt / Ib[1]) * (tb[1] + (Ib[2] - Ib[0]) * wb[2] * wb[0]);
        nw[2] += (dt / Ib[2]) * (tb[2] + (Ib[0] - Ib[1]) * wb[0] * wb[1]);
    }
    body2space(w, nw, q);
}
static void omega_ab2(double *w, const double *tq, const double *tq_old, const double *Ib, const quaternion *q,
                      const quaternion *q_old, int free_, double hdt) {
    double wb[3], tb[3], tbo[3];
    space2body(wb, w, q);
    space2body(tb, tq, q);
    space2body(tbo, tq_old, q_old);
    double nw[3] = {wb[0], wb[1], wb[2]};
    if (free_) {
        nw[0] += (hdt / Ib[0]) * ((tb[0] + tbo[0]) + 2.0 * (Ib[1] - Ib[2]) * wb[1] * wb[2]);
        nw[1] += (hdt / Ib[1]) * ((tb[1] + tbo[1]) + 2.0 * (Ib[2] - Ib[0]) * wb[2] * wb[0]);
        nw[2] += (hdt / Ib[2]) * ((tb[2] + tbo[2]) + 2.0 * (Ib[0] - Ib[1]) * wb[0] * wb[1]);
    }
    body2space(w, nw, q);
}

/* set_Particle_Velocities (rigid.h:218-229): beads move with their body */
static void rigid_bead_velocities(kr_run *r) {
    for (int n = 0; n < r->np; n++) {
        const kr_body *b = &r->body[r->bid[n]];
        double         wxr[3];
        v_cross(wxr, b->wG, r->GR[n], 1.0);
        for (int d = 0; d < 3; d++) {
            r->p[n].omega[d] = b->wG[d];
            r->p[n].v[d]     = b->vG[d] + wxr[d];
        }
    }
}

/* solver_Rigid_Position (rigid.h:333-392), update_Orientation (271-290;
 * the space-frame moments it also rotates feed only the thermostat and
 * the stress, neither of which a rigid run here computes) and
 * update_Particle_Configuration (rigid.h:159-178) */
static void rigid_position(kr_run *r, int ab2) {
    for (int k = 0; k < r->nb; k++) {
        kr_body *b = &r->body[k];
        for (int d = 0; d < 3; d++) {
            b->xGp[d]      = b->xG[d];
            double delta_x = ab2 ? r->jikan.hdt_md * (3.0 * b->vG[d] - b->vGo[d]) : r->jikan.dt_md * b->vG[d];
            b->xG[d] += delta_x;
            b->xGn[d] += delta_x;
        }
        pbc3(r, b->xG);
        Particle *first = &r->p[r->bcum[k]];
        orientation_update(r, first, ab2);
        for (int n = r->bcum[k] + 1; n < r->bcum[k + 1]; n++) {
            r->p[n].q_old = first->q_old;
            r->p[n].q     = first->q;
        }
        for (int n = r->bcum[k]; n < r->bcum[k + 1]; n++) body2space(r->GR[n], r->GRb[n], &first->q);
    }
    for (int n = 0; n < r->np; n++) {
        const kr_body *b = &r->body[r->bid[n]];
        Particle      *p = &r->p[n];
        double         wxr[3];
        v_cross(wxr, b->wG, r->GR[n], 1.0);
        for (int d = 0; d < 3; d++) {
            p->x_previous[d] = p->x[d];
            p->x[d]          = b->xG[d] + r->GR[n][d];
            p->omega_old[d]  = p->omega[d];
            p->v_old[d]      = p->v[d];
            p->omega[d]      = b->wG[d];
            p->v[d]          = b->vG[d] + wxr[d];
        }
        pbc3(r, p->x);
    }
}

/* calc_Rigid_VOGs (rigid.h:467-559), then set_Particle_Velocities */
static void rigid_velocity(kr_run *r, int ab2) {
    for (int k = 0; k < r->nb; k++) {
        kr_body *b = &r->body[k];
        for (int d = 0; d < 3; d++) {
            b->vGo[d] = b->vG[d];
            b->wGo[d] = b->wG[d];
        }
    }
    for (int k = 0; k < r->nb; k++) {
        kr_body        *b     = &r->body[k];
        const Particle *first = &r->p[r->bcum[k]];
        if (!ab2) {
            const double dmy = r->jikan.dt_md / b->M;
            if (b->free_)
                for (int d = 0; d < 3; d++) b->vG[d] += dmy * (b->fG[d] + b->fGr[d]);
            double tq[3] = {b->tG[0] + b->tGr[0], b->tG[1] + b->tGr[1], b->tG[2] + b->tGr[2]};
            omega_euler(b->wG, tq, b->Ib, &first->q, b->free_, r->jikan.dt_md);
        } else {
            const double dmy = r->jikan.hdt_md / b->M;
            if (b->free_)
                for (int d = 0; d < 3; d++) b->vG[d] += dmy * (2.0 * b->fG[d] + b->fGr[d] + b->fGrp[d]);
            double tq[3] = {2.0 * b->tG[0] + b->tGr[0], 2.0 * b->tG[1] + b->tGr[1], 2.0 * b->tG[2] + b->tGr[2]};
            omega_ab2(b->wG, tq, b->tGrp, b->Ib, &first->q, &first->q_old, b->free_, r->jikan.hdt_md);
        }
    }
    for (int k = 0; k < r->nb; k++) {
        kr_body *b = &r->body[k];
        for (int d = 0; d < 3; d++) {
            b->fGp[d]  = b->fG[d];
            b->fGrp[d] = b->fGr[d];
            b->tGp[d]  = b->tG[d];
            b->tGrp[d] = b->tGr[d];
            b->fG[d] = b->fGr[d] = b->tG[d] = b->tGr[d] = 0.0;
        }
    }
    rigid_bead_velocities(r);
}

/* jacobi (matrix_diagonal.cxx:56-172): cyclic Jacobi eigen-decomposition
 * of a symmetric 3x3 matrix (Numerical Recipes' scheme: thresholds for the
 * first three sweeps, negligible elements dropped after four), the
 * eigenvectors returned as rows, ordered by decreasing eigenvalue
 * (jacobi_ordered, 25-53) */
static void jrot(double A[3][3], double s, double tau, int i, int j, int k, int l) {
    double g = A[i][j], h = A[k][l];
    A[i][j]  = g - s * (h + g * tau);
    A[k][l]  = h + s * (g - h * tau);
}
static int jacobi3(const double In[3][3], double V[3][3], double dv[3]) {
    const int n = 3;
    double    A[3][3], b[3], z[3], mval = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) mval = fmax(mval, fabs(In[i][j]));
    const double minv = mval > 0.0 ? 1.0 / mval : 1.0;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            A[i][j] = In[i][j] * minv;
            V[i][j] = 0.0;
        }
        V[i][i] = 1.0;
        b[i] = dv[i] = A[i][i];
        z[i]         = 0.0;
    }
    for (int it = 1; it <= 50; it++) {
        double sm = 0.0;
        for (int ip = 0; ip < n - 1; ip++)
            for (int iq = ip + 1; iq < n; iq++) sm += fabs(A[ip][iq]);
        if (!(sm > TOL_MP || -TOL_MP > sm)) { /* zero_mp(sm) */
            for (int i = 0; i < n - 1; i++)
                for (int j = i + 1; j < n; j++) {
                    double t = V[i][j];
                    V[i][j]  = V[j][i];
                    V[j][i]  = t;
                }
            for (int i = 0; i < n - 1; i++) { /* selection by decreasing eigenvalue, the last of equals */
                int    k   = i;
                double dmy = dv[k];
                for (int j = i; j < n; j++)
                    if (dv[j] >= dmy) {
                        k   = j;
                        dmy = dv[k];
                    }
                if (k != i) {
                    dv[k] = dv[i];
                    dv[i] = dmy;
                    for (int j = 0; j < n; j++) {
                        double t = V[k][j];
                        V[k][j]  = V[i][j];
                        V[i][j]  = t;
                    }
                }
            }
            for (int i = 0; i < n; i++) dv[i] *= mval;
            return 0;
        }
        const double tresh = it < 4 ? 0.2 * sm / (n * n) : 0.0;
        for (int ip = 0; ip < n - 1; ip++)
            for (int iq = ip + 1; iq < n; iq++) {
                const double g = 100.0 * fabs(A[ip][iq]);
                if (it > 4 && g <= 1.0e-17 * fabs(dv[ip]) && g <= 1.0e-17 * fabs(dv[iq])) {
                    A[ip][iq] = 0.0;
                } else if (fabs(A[ip][iq]) > tresh) {
                    double h = dv[iq] - dv[ip], t;
                    if (g <= 1.0e-17 * fabs(h))
                        t = A[ip][iq] / h;
                    else {
                        const double theta = 0.5 * h / A[ip][iq];
                        t                  = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
                        if (theta < 0.0) t = -t;
                    }
                    const double c = 1.0 / sqrt(1 + t * t), sn = t * c, tau = sn / (1.0 + c);
                    h = t * A[ip][iq];
                    z[ip] -= h;
                    z[iq] += h;
                    dv[ip] -= h;
                    dv[iq] += h;
                    A[ip][iq] = 0.0;
                    for (int j = 0; j < ip; j++) jrot(A, sn, tau, j, ip, j, iq);
                    for (int j = ip + 1; j < iq; j++) jrot(A, sn, tau, ip, j, j, iq);
                    for (int j = iq + 1; j < n; j++) jrot(A, sn, tau, ip, j, iq, j);
                    for (int j = 0; j < n; j++) jrot(V, sn, tau, j, ip, j, iq);
                }
            }
        for (int ip = 0; ip < n; ip++) {
            b[ip] += z[ip];
            dv[ip] = b[ip];
            z[ip]  = 0.0;
        }
    }
    return -1;
}

/* M_coordinate_frame (lad3.h:318-354): ex = u/|u|, ey = v orthogonalised,
 * ez = ex x ey, written back into u, v, w */
static void coordinate_frame(double *u, double *v, double *w) {
    double ex[3], ey[3], ez[3], norm = sqrt(SQ(u[0]) + SQ(u[1]) + SQ(u[2]));
    for (int i = 0; i < 3; i++) ex[i] = u[i] / norm;
    double dmy = ex[0] * v[0] + ex[1] * v[1] + ex[2] * v[2];
    for (int i = 0; i < 3; i++) ey[i] = v[i] - dmy * ex[i];
    norm = sqrt(SQ(ey[0]) + SQ(ey[1]) + SQ(ey[2]));
    for (int i = 0; i < 3; i++) ey[i] /= norm;
    ez[0] = ex[1] * ey[2] - ex[2] * ey[1];
    ez[1] = ex[2] * ey[0] - ex[0] * ey[2];
    ez[2] = ex[0] * ey[1] - ex[1] * ey[0];
    norm  = sqrt(SQ(ez[0]) + SQ(ez[1]) + SQ(ez[2]));
    for (int i = 0; i < 3; i++) {
        u[i] = ex[i];
        v[i] = ey[i];
        w[i] = ez[i] / norm;
    }
}

/* rm_rqtn (quaternion.h:491-549): rotation matrix -> quaternion from its
 * largest q_i^2 component */
static void rm_to_qtn(quaternion *q, const double R[3][3]) {
    double qq[4] = {1.0 + R[0][0] + R[1][1] + R[2][2], 1.0 + R[0][0] - R[1][1] - R[2][2],
                    1.0 - R[0][0] + R[1][1] - R[2][2], 1.0 - R[0][0] - R[1][1] + R[2][2]};
    int    i     = 0;
    double mm    = qq[0];
    for (int d = 1; d < 4; d++)
        if (qq[d] > mm) {
            mm = qq[d];
            i  = d;
        }
    double dmy;
    switch (i) {
        case 0:
            qq[0] = sqrt(qq[0]) / 2.0;
            dmy   = 1.0 / (4.0 * qq[0]);
            qq[1] = (R[2][1] - R[1][2]) * dmy;
            qq[2] = (R[0][2] - R[2][0]) * dmy;
            qq[3] = (R[1][0] - R[0][1]) * dmy;
            break;
        case 1:
            qq[1] = sqrt(qq[1]) / 2.0;
            dmy   = 1.0 / (4.0 * qq[1]);
            qq[0] = (R[2][1] - R[1][2]) * dmy;
            qq[2] = (R[1][0] + R[0][1]) * dmy;
            qq[3] = (R[0][2] + R[2][0]) * dmy;
            break;
        case 2:
            qq[2] = sqrt(qq[2]) / 2.0;
            dmy   = 1.0 / (4.0 * qq[2]);
            qq[0] = (R[0][2] - R[2][0]) * dmy;
            qq[1] = (R[1][0] + R[0][1]) * dmy;
            qq[3] = (R[2][1] + R[1][2]) * dmy;
            break;
        default:
            qq[3] = sqrt(qq[3]) / 2.0;
            dmy   = 1.0 / (4.0 * qq[3]);
            qq[0] = (R[1][0] - R[0][1]) * dmy;
            qq[1] = (R[2][0] + R[0][2]) * dmy;
            qq[2] = (R[2][1] + R[1][2]) * dmy;
            break;
    }
    q->s    = qq[0];
    q->v[0] = qq[1];
    q->v[1] = qq[2];
    q->v[2] = qq[3];
}

/* phi-weighted sums of one body over its beads' Sekibun points
 * (Make_phi_rigid_mass / Make_phi_rigid_inertia, make_phi.cxx:589-630,
 * 674-738): mass m, first moment c about the current centre, inertia I */
static void rigid_phi_moments(kr_run *r, int k, double *m, double *c, double J[3][3]) {
    *m = c[0] = c[1] = c[2] = 0.0;
    for (int d = 0; d < 3; d++) J[d][0] = J[d][1] = J[d][2] = 0.0;
    for (int n = r->bcum[k]; n < r->bcum[k + 1]; n++) {
        double xp[3] = {r->p[n].x[0], r->p[n].x[1], r->p[n].x[2]}, residue[3];
        int    x_int[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            const double w = Phi(r, distance(r, x, xp)) / fmax(r->phi_sum[IM(r, r_mesh[0], r_mesh[1], r_mesh[2])], 1.0);
            const double R[3] = {r->GR[n][0] + rv[0], r->GR[n][1] + rv[1], r->GR[n][2] + rv[2]};
            *m += w;
            for (int d = 0; d < 3; d++) c[d] += w * R[d];
            J[0][0] += w * (R[1] * R[1] + R[2] * R[2]);
            J[0][1] += w * (-R[0] * R[1]);
            J[0][2] += w * (-R[0] * R[2]);
            J[1][0] += w * (-R[1] * R[0]);
            J[1][1] += w * (R[0] * R[0] + R[2] * R[2]);
            J[1][2] += w * (-R[1] * R[2]);
            J[2][0] += w * (-R[2] * R[0]);
            J[2][1] += w * (-R[2] * R[1]);
            J[2][2] += w * (R[0] * R[0] + R[1] * R[1]);
        }
    }
}

/* The initial block of a rigid run (init_particle.cxx:455-487): trial
 * centres from the unwrapped bead coordinates, the phi-weighted centre
 * and mass, bead vectors, inertia, principal frame, bead velocities. */
int kr_set_rigid(kr_run *r, int nb, const int *cum, const int *free_flags, const double *v0, const double *w0) {
    if (nb < 1 || !cum || cum[0] != 0 || cum[nb] != r->np || !r->P.rotation || r->kBT > 0. || r->sw_janus_slip ||
        r->le)
        return -1;
    for (int k = 0; k < nb; k++)
        if (cum[k + 1] <= cum[k]) return -1;
    free(r->bcum);
    free(r->bid);
    free(r->body);
    free(r->GR);
    free(r->GRb);
    r->nb   = nb;
    r->bcum = (int *)malloc(sizeof(int) * (size_t)(nb + 1));
    r->bid  = (int *)malloc(sizeof(int) * (size_t)r->np);
    r->body = (kr_body *)calloc((size_t)nb, sizeof(kr_body));
    r->GR   = (double(*)[3])calloc((size_t)r->np, sizeof(double[3]));
    r->GRb  = (double(*)[3])calloc((size_t)r->np, sizeof(double[3]));
    memcpy(r->bcum, cum, sizeof(int) * (size_t)(nb + 1));
    for (int k = 0; k < nb; k++) {
        kr_body *b = &r->body[k];
        b->spec    = r->p[cum[k]].spec;
        b->free_   = free_flags ? (free_flags[k] != 0) : 1;
        for (int n = cum[k]; n < cum[k + 1]; n++) r->bid[n] = k;
        /* init_set_xGs (rigid.h:25-38) on the unwrapped coordinates */
        for (int n = cum[k]; n < cum[k + 1]; n++)
            for (int d = 0; d < 3; d++) b->xG[d] += r->p[n].x_nopbc[d];
        for (int d = 0; d < 3; d++) b->xG[d] /= (double)(cum[k + 1] - cum[k]);
        /* init_set_PBC (rigid.h:58-80) */
        for (int d = 0; d < 3; d++) b->xGn[d] = b->xGp[d] = b->xG[d];
        for (int n = cum[k]; n < cum[k + 1]; n++)
            for (int d = 0; d < 3; d++) r->p[n].x_previous[d] = r->p[n].x_nopbc[d];
        pbc3(r, b->xG);
    }
    /* init_set_GRvecs (rigid.h:44-54) */
    for (int n = 0; n < r->np; n++)
        for (int d = 0; d < 3; d++) r->GR[n][d] = r->p[n].x_nopbc[d] - r->body[r->bid[n]].xGn[d];
    reset_field(r, r->phi);
    reset_field(r, r->phi_sum);
    make_phi_particle_sum(r, r->phi, r->phi_sum);
    for (int k = 0; k < nb; k++) { /* Make_phi_rigid_mass */
        kr_body     *b = &r->body[k];
        double       m, c[3], J[3][3];
        const double rho_p = r->spec[b->spec].mass_ratio * r->RHO;
        rigid_phi_moments(r, k, &m, c, J);
        b->M = m * r->DX3 * rho_p;
        for (int d = 0; d < 3; d++) b->xGn[d] += c[d] / m;
        for (int d = 0; d < 3; d++) b->xG[d] = b->xGn[d];
        pbc3(r, b->xG);
    }
    for (int n = 0; n < r->np; n++)
        for (int d = 0; d < 3; d++) r->GR[n][d] = r->p[n].x_nopbc[d] - r->body[r->bid[n]].xGn[d];
    for (int k = 0; k < nb; k++) { /* Make_phi_rigid_inertia, init_Rigid_Coordinates (rigid.h:121-156) */
        kr_body     *b = &r->body[k];
        double       m, c[3], J[3][3], V[3][3], ev[3], R[3][3];
        const double rho_p = r->spec[b->spec].mass_ratio * r->RHO;
        rigid_phi_moments(r, k, &m, c, J);
        for (int d = 0; d < 3; d++)
            for (int e = 0; e < 3; e++) b->Mom[3 * d + e] = J[d][e] = J[d][e] * r->DX3 * rho_p;
        if (jacobi3(J, V, ev)) return -1;
        coordinate_frame(V[2], V[0], V[1]);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[j][i] = V[i][j];
        quaternion q;
        rm_to_qtn(&q, R);
        for (int n = cum[k]; n < cum[k + 1]; n++) r->p[n].q = q;
        /* Rigid_Moments_body = R^T I R (rigid_body.cxx:26-60); its diagonal */
        for (int d = 0; d < 3; d++) {
            double col[3] = {J[0][d], J[1][d], J[2][d]}, rc[3];
            space2body(rc, col, &q);
            for (int e = 0; e < 3; e++) R[e][d] = rc[e]; /* R := R^T I, by columns */
        }
        for (int d = 0; d < 3; d++) {
            double row[3] = {R[d][0], R[d][1], R[d][2]}, rr[3];
            space2body(rr, row, &q);
            b->Ib[d] = rr[d];
        }
        for (int n = cum[k]; n < cum[k + 1]; n++) space2body(r->GRb[n], r->GR[n], &q);
    }
    /* init_set_vGs (rigid.h:230-240): the bodies' initial velocities (space
     * frame; the reference starts them at rest) carried to the beads */
    for (int k = 0; k < nb; k++)
        for (int d = 0; d < 3; d++) {
            r->body[k].vG[d] = v0 ? v0[3 * k + d] : 0.0;
            r->body[k].wG[d] = w0 ? w0[3 * k + d] : 0.0;
        }
    rigid_bead_velocities(r);
    return 0;
}

/* per body: xG, xG_nopbc, vG, omegaG, the previous step's hydrodynamic and
 * LJ + gravity force and torque, mass, principal moments, orientation */
void kr_get_rigid(kr_run *r, double *out) {
    for (int k = 0; k < r->nb; k++) {
        const kr_body  *b = &r->body[k];
        const Particle *f = &r->p[r->bcum[k]];
        double         *o = out + (size_t)KH_RIGID_WORDS * k;
        for (int d = 0; d < 3; d++) {
            o[d]      = b->xG[d];
            o[3 + d]  = b->xGn[d];
            o[6 + d]  = b->vG[d];
            o[9 + d]  = b->wG[d];
            o[12 + d] = b->fGp[d];
            o[15 + d] = b->tGp[d];
            o[18 + d] = b->fGrp[d];
            o[21 + d] = b->tGrp[d];
            o[25 + d] = b->Ib[d];
        }
        o[24] = b->M;
        o[28] = f->q.s;
        o[29] = f->q.v[0];
        o[30] = f->q.v[1];
        o[31] = f->q.v[2];
    }
}

/* ------------------------------------------------------------------ */
/* thermal noise: Philox4x32-10 (Salmon et al., SC'11) + Box-Muller     */
/* ------------------------------------------------------------------ */
void kr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0          = n0;
        c1          = n1;
        c2          = n2;
        c3          = n3;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static double u53(uint32_t hi, uint32_t lo) { /* open interval (0,1) */
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6) + 0.5) * (1.0 / 9007199254740992.0);
}

void kr_noise_normals(uint64_t seed, uint32_t stream, int particle, int ts, double out6[6]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int c = 0; c < 3; c++) {
        uint32_t ctr[4] = {(uint32_t)particle, (uint32_t)ts, (uint32_t)c, stream}, o[4];
        kr_philox4x32_10(ctr, key, o);
        double u0 = u53(o[0], o[1]), u1 = u53(o[2], o[3]);
        double rad = sqrt(-2.0 * log(u0));
        double ang = 2.0 * KR_PI * u1;
        out6[2 * c]     = rad * cos(ang);
        out6[2 * c + 1] = rad * sin(ang);
    }
}

/* Add_random_force_thermostat (fluct.cxx:10-43), non-rigid branch */
static void thermostat(kr_run *r) {
    const double Zeta_drag     = 6. * KR_PI * r->ETA * r->RADIUS;
    const double Zeta_drag_rot = 8. * KR_PI * r->ETA * r->RADIUS * r->RADIUS * r->RADIUS;
    const double sdv_v         = sqrt(Zeta_drag * r->jikan.dt_md * r->kBT * r->P.alpha_v);
    const double sdv_omega     = sqrt(Zeta_drag_rot * r->jikan.dt_md * r->kBT * r->P.alpha_o);
    const double niv = kT_snap_v * sdv_v, nio = kT_snap_o * sdv_omega;
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        if (is_obstacle(r, p)) continue;
        double dmy[6];
        kr_noise_normals(r->P.seed, r->P.stream_id, n, r->jikan.ts, dmy);
        double imass = r->IMASS[p->spec], imoi = r->IMOI[p->spec];
        for (int d = 0; d < 3; d++) {
            p->v[d] += dmy[d] * niv * imass;
            if (r->P.rotation) p->omega[d] += dmy[d + 3] * nio * imoi;
        }
    }
}

/* ------------------------------------------------------------------ */
/* squirmer slip: operate_surface.cxx, make_phi.h:282-376, md_force.cxx:325-382 */
/* ------------------------------------------------------------------ */
/* Squirmer_coord (make_phi.h:282-376) */
static void squirmer_coord_q(int ax, const quaternion *qp, const double *x, double *nr, double *th, double *ph,
                             double *norm_r, double *theta_angle, double *phi_angle) {
    static const double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0}, ez[3] = {0, 0, 1};
    const double       *e1, *e2, *e3;
    if (ax == KH_JAX_Z) {
        e1 = ex;
        e2 = ey;
        e3 = ez;
    } else if (ax == KH_JAX_Y) {
        e1 = ez;
        e2 = ex;
        e3 = ey;
    } else {
        e1 = ey;
        e2 = ez;
        e3 = ex;
    }
    space2body(nr, x, qp);
    *norm_r      = sqrt(SQ(nr[0]) + SQ(nr[1]) + SQ(nr[2]));
    double dnorm = 1.0 / *norm_r;
    for (int d = 0; d < 3; d++) nr[d] *= dnorm;
    v_cross(ph, e3, nr, 1.0);
    double dot_e3_r = e3[0] * nr[0] + e3[1] * nr[1] + e3[2] * nr[2];
    dnorm           = sqrt(SQ(ph[0]) + SQ(ph[1]) + SQ(ph[2]));
    if (((dnorm > TOL_MP) || (-TOL_MP > dnorm)) && less_than_mp(fabs(dot_e3_r), 1.0)) {
        dnorm = 1.0 / dnorm;
        for (int d = 0; d < 3; d++) ph[d] *= dnorm;
        *theta_angle = acos(dot_e3_r);
        v_cross(th, ph, nr, 1.0);
        dnorm = 1.0 / sqrt(SQ(th[0]) + SQ(th[1]) + SQ(th[2]));
        for (int d = 0; d < 3; d++) th[d] *= dnorm;
        double r12[3];
        for (int d = 0; d < 3; d++) r12[d] = nr[d] - e3[d] * dot_e3_r;
        double d1  = e1[0] * r12[0] + e1[1] * r12[1] + e1[2] * r12[2];
        double d2  = e2[0] * r12[0] + e2[1] * r12[1] + e2[2] * r12[2];
        *phi_angle = atan2(d2, d1);
        double tmp[3];
        body2space(tmp, nr, qp);
        memcpy(nr, tmp, sizeof tmp);
        body2space(tmp, th, qp);
        memcpy(th, tmp, sizeof tmp);
        body2space(tmp, ph, qp);
        memcpy(ph, tmp, sizeof tmp);
    } else {
        *norm_r = sqrt(SQ(x[0]) + SQ(x[1]) + SQ(x[2]));
        dnorm   = 1.0 / *norm_r;
        for (int d = 0; d < 3; d++) {
            nr[d] = x[d] * dnorm;
            ph[d] = th[d] = 0.0;
        }
        *theta_angle = (dot_e3_r > 0.0) ? 0.0 : KR_PI;
        *phi_angle   = 0.0;
    }
}

static void squirmer_coord(const kr_run *r, const double *x, double *nr, double *th, double *ph, double *norm_r,
                           double *theta_angle, double *phi_angle, const Particle *p) {
    squirmer_coord_q(r->spec[p->spec].janus_axis, &p->q, x, nr, th, ph, norm_r, theta_angle, phi_angle);
}

void kr_squirmer_coord(const double q4[4], int axis, const double x[3], double nr[3], double nth[3], double nph[3],
                       double *theta, double *phi) {
    quaternion q;
    q.s = q4[0];
    for (int d = 0; d < 3; d++) q.v[d] = q4[1 + d];
    double norm_r;
    squirmer_coord_q(axis, &q, x, nr, nth, nph, &norm_r, theta, phi);
}

/* Make_particle_momentum_factor (operate_surface.cxx:12-153) */
static void make_particle_momentum_factor(kr_run *r, double **u) {
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p  = &r->p[n];
        int       ps = p->spec;
        double    xp[3], M0 = 0, SM0 = 0, M1[3] = {0}, SM1[3] = {0}, M2[3][3] = {{0}}, SM2[3][3] = {{0}};
        double    dv_s[3] = {0}, dw_s[3] = {0};
        for (int d = 0; d < 3; d++) xp[d] = p->x[d];
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3], uf[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            size_t im = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            for (int d = 0; d < 3; d++) {
                x[d]  = r_mesh[d] * r->DX;
                uf[d] = u[d][im];
            }
            double dmy_r   = distance(r, x, xp);
            double dmy_phi = Phi(r, dmy_r);
            double dmy_xi  = fabs(dmy_r - r->RADIUS);
            M0 += dmy_phi;
            for (int d = 0; d < 3; d++) {
                M1[d] += dmy_phi * rv[d];
                M2[d][d] += dmy_phi * dmy_r * dmy_r;
                for (int l = 0; l < 3; l++) M2[d][l] -= dmy_phi * rv[d] * rv[l];
            }
            if (r->spec[ps].propulsion == KH_JP_SLIP && less_than_mp(dmy_xi, r->HXI)) {
                double slip_vel = r->spec[ps].slip_vel,
                       slip_mode = r->slip_mode_half[ps] * (g_mut == KR_MUT_SLIP_B2_FULL ? 2.0 : 1.0);
                double C1 = r->spec[ps].rotlet_c1 * (g_mut == KR_MUT_SLIP_ROTLET ? 2.0 : 1.0), C2 = r->spec[ps].rotlet_c2;
                double dmy_phi_s = (1.0 - dmy_phi) * DPhi_compact_sin_norm(r, dmy_r);
                SM0 += dmy_phi_s;
                double n_r[3], n_th[3], n_ta[3], sr, th, ta;
                squirmer_coord(r, rv, n_r, n_th, n_ta, &sr, &th, &ta, p);
                double slip_mag = slip_vel * (sin(th) + slip_mode * sin(2.0 * th));
                double slip_az  = C1 * sin(th) + C2 * 1.5 * sin(2.0 * th);
                double us[3], rxus[3];
                for (int d = 0; d < 3; d++) us[d] = n_th[d] * slip_mag + n_ta[d] * slip_az - uf[d];
                v_cross(rxus, rv, us, 1.0);
                for (int d = 0; d < 3; d++) {
                    dv_s[d] += dmy_phi_s * us[d];
                    dw_s[d] += dmy_phi_s * rxus[d];
                    SM1[d] += dmy_phi_s * rv[d];
                    SM2[d][d] += dmy_phi_s * dmy_r * dmy_r;
                    for (int l = 0; l < 3; l++) SM2[d][l] -= dmy_phi_s * rv[d] * rv[l];
                }
            }
        }
        p->mass         = M0;
        p->surface_mass = SM0;
        for (int d = 0; d < 3; d++) {
            p->mass_center[d]         = M1[d];
            p->surface_mass_center[d] = SM1[d];
            p->surface_dv[d]          = dv_s[d];
            p->surface_dw[d]          = dw_s[d];
            for (int l = 0; l < 3; l++) {
                p->inertia[d][l]         = M2[d][l];
                p->surface_inertia[d][l] = SM2[d][l];
            }
        }
    }
}

/* M_det / M_inv (lad3.h:202-233) */
static void m_inv_inplace(double B[3][3]) {
    double A[3][3];
    memcpy(A, B, sizeof A);
    double idetA = -A[0][2] * A[1][1] * A[2][0] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
                   A[0][0] * A[1][2] * A[2][1] - A[0][1] * A[1][0] * A[2][2] + A[0][0] * A[1][1] * A[2][2];
    idetA   = 1.0 / idetA;
    B[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * idetA;
    B[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * idetA;
    B[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * idetA;
    B[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * idetA;
    B[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * idetA;
    B[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * idetA;
    B[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * idetA;
    B[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * idetA;
    B[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * idetA;
}

/* slip_droplet (operate_surface.cxx:155-198) */
static void slip_droplet(double *vp, double *wp, double *delta_v, double *delta_w, const Particle *p) {
    double LL[3][3], dP[3], dL[3], L0[3], mc[3], imass, mc2;
    for (int d = 0; d < 3; d++) {
        vp[d] = p->v_slip[d];
        wp[d] = p->omega_slip[d];
        mc[d] = p->mass_center[d];
        LL[d][0] = LL[d][1] = LL[d][2] = 0.0;
    }
    imass = 1.0 / p->mass;
    double w_x_r[3], r_x_v[3];
    v_cross(w_x_r, wp, p->surface_mass_center, 1.0);
    v_cross(r_x_v, p->surface_mass_center, vp, 1.0);
    for (int d = 0; d < 3; d++) {
        dP[d] = -(p->surface_mass * vp[d] + w_x_r[d] + p->surface_dv[d]);
        dL[d] = -(r_x_v[d] + p->surface_inertia[d][0] * wp[0] + p->surface_inertia[d][1] * wp[1] +
                  p->surface_inertia[d][2] * wp[2] + p->surface_dw[d]);
    }
    v_cross(L0, mc, dP, imass);
    mc2 = -(mc[0] * mc[0] + mc[1] * mc[1] + mc[2] * mc[2]) * imass;
    for (int d = 0; d < 3; d++) {
        L0[d]    = dL[d] - L0[d];
        LL[d][d] = mc2;
        for (int l = 0; l < 3; l++) LL[d][l] += (p->inertia[d][l] + imass * mc[d] * mc[l]);
    }
    m_inv_inplace(LL);
    for (int i = 0; i < 3; i++) { /* M_v_prod (lad3.h:457-466) */
        double dmy = 0.0;
        for (int j = 0; j < 3; j++) dmy += LL[i][j] * L0[j];
        delta_w[i] = dmy;
    }
    v_cross(L0, delta_w, mc, 1.0);
    for (int d = 0; d < 3; d++) delta_v[d] = imass * (dP[d] - L0[d]);
}

/* Make_force_u_slip_particle (operate_surface.cxx:200-363), momentum check omitted */
static void make_force_u_slip_particle(kr_run *r, double **up, double **u) {
    for (int n = 0; n < r->np; n++) { /* serial: scatter order fixed */
        Particle *p  = &r->p[n];
        int       ps = p->spec;
        for (int d = 0; d < 3; d++) {
            p->f_slip[d]      = 0.0;
            p->torque_slip[d] = 0.0;
        }
        if (r->spec[ps].propulsion != KH_JP_SLIP) continue;
        double slip_vel = r->spec[ps].slip_vel,
                       slip_mode = r->slip_mode_half[ps] * (g_mut == KR_MUT_SLIP_B2_FULL ? 2.0 : 1.0);
        double C1 = r->spec[ps].rotlet_c1 * (g_mut == KR_MUT_SLIP_ROTLET ? 2.0 : 1.0), C2 = r->spec[ps].rotlet_c2;
        double vp[3], om[3], dv[3], dw[3], xp[3];
        slip_droplet(vp, om, dv, dw, p);
        for (int d = 0; d < 3; d++) xp[d] = p->x[d];
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3], uf[3], fv[3], dvrot[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            size_t im = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            for (int d = 0; d < 3; d++) {
                uf[d] = u[d][im];
                x[d]  = r_mesh[d] * r->DX;
            }
            double dmy_r   = distance(r, x, xp);
            double dmy_phi = Phi(r, dmy_r);
            double dmy_xi  = fabs(dmy_r - r->RADIUS);
            angular2v(r, dw, rv, dvrot);
            for (int d = 0; d < 3; d++) {
                fv[d] = dmy_phi * (dv[d] + dvrot[d]);
                up[d][im] += fv[d];
            }
            double dmy_phi_s = (1.0 - dmy_phi) * DPhi_compact_sin_norm(r, dmy_r);
            if (less_than_mp(dmy_xi, r->HXI)) {
                double v_rot[3], n_r[3], n_th[3], n_ta[3], sr, th, ta;
                angular2v(r, om, rv, v_rot);
                squirmer_coord(r, rv, n_r, n_th, n_ta, &sr, &th, &ta, p);
                double vs  = slip_vel * (sin(th) + slip_mode * sin(2.0 * th));
                double vs2 = C1 * sin(th) + C2 * 1.5 * sin(2.0 * th);
                for (int d = 0; d < 3; d++)
                    fv[d] = dmy_phi_s * (vp[d] + v_rot[d] + n_th[d] * vs + n_ta[d] * vs2 - uf[d]);
                for (int d = 0; d < 3; d++) up[d][im] += fv[d];
            }
        }
    }
}

/* Calc_f_slip_correct_precision (md_force.cxx:325-382) */
static void calc_f_slip(kr_run *r, double **u) {
    const double dmy = (g_mut == KR_MUT_SLIP_FORCE_SIGN ? -1.0 : 1.0) * (r->DX3 * r->RHO) / r->jikan.dt_fluid;
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        double    xp[3], force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
        for (int d = 0; d < 3; d++) xp[d] = p->x[d];
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3], fp[3];
            relative_coord(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            double dmyR    = distance(r, x, xp);
            double dmy_phi = Phi(r, dmyR);
            size_t im      = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            for (int d = 0; d < 3; d++) {
                fp[d] = u[d][im] * dmy_phi;
                force[d] += fp[d];
            }
            torque[0] += (rv[1] * fp[2] - rv[2] * fp[1]);
            torque[1] += (rv[2] * fp[0] - rv[0] * fp[2]);
            torque[2] += (rv[0] * fp[1] - rv[1] * fp[0]);
        }
        for (int d = 0; d < 3; d++) {
            p->f_slip[d]      = (dmy * force[d]);
            p->torque_slip[d] = (dmy * torque[d]);
        }
    }
}

/* MD_solver_velocity_slip_Euler / _AB2 (particle_solver.cxx:144-207) */
static void velocity_slip(kr_run *r, int ab2) {
    for (int n = 0; n < r->np; n++) {
        Particle *p    = &r->p[n];
        int       s    = p->spec;
        double    h    = ab2 ? r->jikan.hdt_md : r->jikan.dt_md;
        double    dmy  = h * r->IMASS[s];
        double    dmyr = h * r->IMOI[s];
        if (!is_obstacle(r, p)) {
            double sf[3], st[3];
            self_propulsion(r, p, sf, st);
            for (int d = 0; d < 3; d++) {
                p->v_old[d] = p->v[d];
                if (!ab2)
                    p->v[d] += (dmy * (p->f_hydro[d] + p->f_slip[d] + p->fr[d] + sf[d]));
                else
                    p->v[d] += dmy * (2.0 * (p->f_hydro[d] + p->f_slip[d] + sf[d]) + p->fr[d] + p->fr_previous[d]);
                p->omega_old[d] = p->omega[d];
                if (!ab2)
                    p->omega[d] += (dmyr * (p->torque_hydro[d] + p->torque_slip[d] + p->torque_r[d] + st[d]));
                else
                    p->omega[d] += dmyr * (2.0 * (p->torque_hydro[d] + p->torque_slip[d] + st[d]) + p->torque_r[d] +
                                           p->torque_r_previous[d]);
            }
        } else {
            for (int d = 0; d < 3; d++) {
                p->v_old[d]     = p->v[d];
                p->v[d]         = 0.0;
                p->omega_old[d] = p->omega[d];
                p->omega[d]     = 0.0;
            }
        }
    }
}

/* Slip_particle_convergence (operate_surface.h:72-100) */
static double slip_convergence(kr_run *r) {
    double eps = 0.0;
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        double    dv = 0, dw = 0, nv = 0, nw = 0;
        if (r->spec[p->spec].propulsion == KH_JP_SLIP) {
            for (int d = 0; d < 3; d++) {
                nv += p->v[d] * p->v[d];
                nw += p->omega[d] * p->omega[d];
                dv += (p->v_slip[d] - p->v[d]) * (p->v_slip[d] - p->v[d]);
                dw += (p->omega_slip[d] - p->omega[d]) * (p->omega_slip[d] - p->omega[d]);
            }
            if (nv > TOL_MP) dv /= nv;
            if (nw > TOL_MP) dw /= nw;
        }
        eps = fmax(eps, dv + dw);
    }
    return sqrt(eps);
}

/* Update_slip_particle_velocity (operate_surface.h:101-121) */
static void update_slip_velocity(kr_run *r, int iter) {
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        for (int d = 0; d < 3; d++) {
            if (iter == 0) {
                p->v_slip[d]     = p->v[d];
                p->omega_slip[d] = p->omega[d];
            } else {
                const double w   = g_mut == KR_MUT_SLIP_RELAX ? 0.5 : 0.7;
                p->v_slip[d]     = (1.0 - w) * p->v_slip[d] + w * p->v[d];
                p->omega_slip[d] = (1.0 - w) * p->omega_slip[d] + w * p->omega[d];
            }
        }
    }
}

/* the slip block of Time_evolution_hydro (sp_3d_ns.cxx:121-161) */
static void slip_iteration(kr_run *r) {
    int converge = 0, iter = 0;
    make_particle_momentum_factor(r, r->u);
    update_slip_velocity(r, 0);
    while (!converge) {
        for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
        make_force_u_slip_particle(r, r->up, r->u);
        solenoidal_u(r, r->up);
        calc_f_slip(r, r->up);
        add_f_particle(r, r->up, r->u); /* up += u */
        if (iter == 0) force(r);        /* MD_solver_velocity_slip_iter(start_iter) */
        velocity_slip(r, r->jikan.ts != 0);
        iter++;
        if (slip_convergence(r) < r->P.slip_tol || iter == r->P.slip_iter_max) {
            converge = 1;
            reset_forces(r); /* end_iter */
        } else {
            update_slip_velocity(r, iter);
            for (int n = 0; n < r->np && g_mut != KR_MUT_SLIP_NO_RESET; n++) /* reset_iter */
                for (int d = 0; d < 3; d++) {
                    r->p[n].v[d]     = r->p[n].v_old[d];
                    r->p[n].omega[d] = r->p[n].omega_old[d];
                }
        }
    }
    for (int d = 0; d < 3 && g_mut != KR_MUT_SLIP_FINAL_USTAR; d++) { /* Swap_mem(u, up) */
        double *t = r->u[d];
        r->u[d]   = r->up[d];
        r->up[d]  = t;
    }
}

/* ------------------------------------------------------------------ */
/* Lees-Edwards sheared step (SW_EQ == Shear_Navier_Stokes_Lees_Edwards) */
/* ------------------------------------------------------------------ */
/* The fluid lives on an oblique grid that follows the mean shear
 * u_x = shear_rate_eff * (y - LY/2): velocities are contravariant
 * components on that grid, and degree_oblique is the accumulated strain,
 * wrapped to (-XYaspect, XYaspect) by a remap of the grid. Spherical,
 * non-rigid particles; linear interpolation between the grids
 * (SW_OBL_INT == linear_int). The shear-stress diagnostics (Hydro_force,
 * Calc_Reynolds_shear_stress, dev_shear_stress_*) are not computed: they
 * feed output only and change no state. */

/* contra2co_single / co2contra_single (fft_wrapper.h:611-633) */
static void contra2co_single(double g, double *v) {
    double a = v[0], b = v[1];
    v[0]     = a + g * b;
    v[1]     = g * a + (1. + g * g) * b;
}

/* Zeta_k2u_k_OBL (fft_wrapper.cxx:690-768): contravariant u from zeta */
static void zeta_k2u_k_obl(kr_run *r, double **zeta, const double *uk_dc, double **u) {
    const double g = r->degree_oblique;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int kk = 0; kk < r->HNZ_; kk++) {
                size_t im = IM(r, i, j, 2 * kk);
                double kx = r->WAVE[0] * r->KX[im], ky = r->WAVE[1] * r->KY[im], kz = r->WAVE[2] * r->KZ[im];
                double ik2 = r->IK2[im];
                double d1r = zeta[0][im], d1i = zeta[0][im + 1], d2r = zeta[1][im], d2i = zeta[1][im + 1];
                double ore[3], oim[3];
                if (r->KX[im] != 0) {
                    ore[0] = -(1. / kx) * (ky * d1r + kz * d2r);
                    oim[0] = -(1. / kx) * (ky * d1i + kz * d2i);
                    ore[1] = d1r;
                    oim[1] = d1i;
                    ore[2] = d2r;
                    oim[2] = d2i;
                } else if (r->KZ[im] != 0) {
                    double dmy = -(ky / kz);
                    ore[0]     = d2r;
                    oim[0]     = d2i;
                    ore[1]     = d1r;
                    oim[1]     = d1i;
                    ore[2]     = dmy * d1r;
                    oim[2]     = dmy * d1i;
                } else {
                    ore[0] = d1r;
                    oim[0] = d1i;
                    ore[1] = 0.0;
                    oim[1] = 0.0;
                    ore[2] = d2r;
                    oim[2] = d2i;
                }
                contra2co_single(g, ore);
                contra2co_single(g, oim);
                kx *= ik2;
                ky *= ik2;
                kz *= ik2;
                u[0][im]     = -(ky * oim[2] - kz * oim[1]);
                u[0][im + 1] = (ky * ore[2] - kz * ore[1]);
                u[1][im]     = -(kz * oim[0] - kx * oim[2]);
                u[1][im + 1] = (kz * ore[0] - kx * ore[2]);
                u[2][im]     = -(kx * oim[1] - ky * oim[0]);
                u[2][im + 1] = (kx * ore[1] - ky * ore[0]);
            }
    u[0][0] = uk_dc[0];
    u[1][0] = uk_dc[1];
    u[2][0] = uk_dc[2];
}

/* Zeta_k2omega_k_OBL (fft_wrapper.cxx:820-865): contravariant omega */
static void zeta_k2omega_k_obl(kr_run *r, double **zeta, double **omega) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                double kx = r->WAVE[0] * r->KX[im], ky = r->WAVE[1] * r->KY[im], kz = r->WAVE[2] * r->KZ[im];
                double d1 = zeta[0][im], d2 = zeta[1][im];
                if (r->KX[im] != 0) {
                    omega[0][im] = -(1. / kx) * (ky * d1 + kz * d2);
                    omega[1][im] = d1;
                    omega[2][im] = d2;
                } else if (r->KZ[im] != 0) {
                    omega[0][im] = d2;
                    omega[1][im] = d1;
                    omega[2][im] = -(ky / kz) * d1;
                } else {
                    omega[0][im] = d1;
                    omega[1][im] = 0.0;
                    omega[2][im] = d2;
                }
            }
    for (int d = 0; d < 3; d++) omega[d][0] = 0.0;
}

/* U_k2rotation_k (fft_wrapper.cxx:893-935): u <- i k x u */
static void u_k2rotation_k(kr_run *r, double **u) {
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int kk = 0; kk < r->HNZ_; kk++) {
                size_t im    = IM(r, i, j, 2 * kk);
                double ks[3] = {r->KX[im] * r->WAVE[0], r->KY[im] * r->WAVE[1], r->KZ[im] * r->WAVE[2]};
                double ur[3], ui[3];
                for (int d = 0; d < 3; d++) {
                    ur[d] = u[d][im];
                    ui[d] = u[d][im + 1];
                }
                u[0][im]     = -(ks[1] * ui[2] - ks[2] * ui[1]);
                u[0][im + 1] = (ks[1] * ur[2] - ks[2] * ur[1]);
                u[1][im]     = -(ks[2] * ui[0] - ks[0] * ui[2]);
                u[1][im + 1] = (ks[2] * ur[0] - ks[0] * ur[2]);
                u[2][im]     = -(ks[0] * ui[1] - ks[1] * ui[0]);
                u[2][im + 1] = (ks[0] * ur[1] - ks[1] * ur[0]);
            }
}

/* Zeta_k2advection_k_OBL (operate_omega.cxx:100-155) with Zeta_k2u_cpuky
 * (operate_omega.h:153-157), Zeta_k2omega_OBL (191-197) and
 * Omega_k2zeta_k_OBL (fft_wrapper.cxx:265-289) */
static void zeta_k2advection_k_obl(kr_run *r, double **zeta, double *uk_dc, double **adv) {
    truncate_two_third(r, zeta[0]);
    truncate_two_third(r, zeta[1]);
    double **u = r->u, *w1 = r->w1, **w3 = r->w3;
    zeta_k2u_k_obl(r, zeta, uk_dc, u);
    memcpy(w1, u[1], sizeof(double) * r->M);
    for (int d = 0; d < 3; d++) A_k2a(r, u[d]);
    zeta_k2omega_k_obl(r, zeta, w3);
    for (int d = 0; d < 3; d++) A_k2a(r, w3[d]);
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                double u1 = u[0][im], u2 = u[1][im], u3 = u[2][im];
                double o1 = w3[0][im], o2 = w3[1][im], o3 = w3[2][im];
                u[0][im]  = u2 * o3 - u3 * o2;
                u[1][im]  = u3 * o1 - u1 * o3;
                u[2][im]  = u1 * o2 - u2 * o1;
            }
    for (int d = 0; d < 3; d++) A2a_k(r, u[d]);
    const double s = r->shear_rate_eff, g = r->degree_oblique;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                u[0][im] -= 2. * s * w1[im];
                u[1][im] -= 2. * s * g * w1[im];
            }
    u_k2rotation_k(r, u);
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                if (r->KX[im] != 0) {
                    adv[0][im] = u[1][im];
                    adv[1][im] = u[2][im];
                } else if (r->KZ[im] != 0) {
                    adv[0][im] = u[1][im];
                    adv[1][im] = u[0][im];
                } else {
                    adv[0][im] = u[0][im];
                    adv[1][im] = u[2][im];
                }
            }
}

/* NS_solver_slavedEuler_Shear_OBL (fluid_solver.cxx:411-437) */
static void ns_solver_obl(kr_run *r) {
    zeta_k2advection_k_obl(r, r->zeta, r->uk_dc, r->f_ns0);
    slaved_euler_update(r);
}

/* Update_K2_OBL (fluid_solver.h:342-360) */
static void update_k2_obl(kr_run *r) {
    const double g = r->degree_oblique;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                r->K2[im] = SQ(r->WAVE[0] * r->KX[im]) + SQ(r->WAVE[1] * r->KY[im] - r->WAVE[0] * r->KX[im] * g) +
                            SQ(r->WAVE[2] * r->KZ[im]);
                r->IK2[im] = (r->K2[im] > 0.0) ? 1.0 / r->K2[im] : 0.0;
            }
}

/* Reset_U_OBL (fluid_solver.h:312-340): the oblique grid at strain
 * +-XYaspect re-expressed on the grid at strain 0 (flg 0: from +XYaspect). */
static void reset_u_obl(kr_run *r, double **u, double **ucp, int flg) {
    const double delta = (flg == 0 ? r->XYaspect : -r->XYaspect);
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++) {
            double sign = j - r->NY / 2;
            if (!(sign == 0)) sign = sign / fabs(sign);
            int i_oblique = (int)(sign * (j - r->NY / 2)) * sign;
            i_oblique     = (int)fmod(i + delta * i_oblique + 4. * r->NX, r->NX);
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k), io = IM(r, i_oblique, j, k);
                u[0][io]  = ucp[0][im] + delta * ucp[1][im];
                u[1][io]  = ucp[1][im];
                u[2][io]  = ucp[2][im];
            }
        }
}

/* obl2orth (fft_wrapper.h:276-288) */
static void obl2orth(const kr_run *r, int j, int i, int *i_plus, int *i_oblique, double *alpha, double *beta) {
    const double g       = r->degree_oblique;
    int          delta_j = j - r->NY / 2;
    double       sign    = (g >= 0.0 ? (double)delta_j : -(double)delta_j);
    if (!(delta_j == 0)) sign = sign / fabs(sign);
    int io     = (int)((int)(sign * g * delta_j) * sign + sign);
    *alpha     = (io - g * delta_j) * sign;
    *beta      = 1.0 - *alpha;
    *i_oblique = (int)fmod(i + io + 4.0 * r->NX, r->NX);
    *i_plus    = (int)fmod(i + sign + 2.0 * r->NX, r->NX);
}
/* orth2obl (fft_wrapper.h:262-274) */
static void orth2obl(const kr_run *r, int j, int i, int *i_oblique, int *i_oblique_plus, double *alpha, double *beta) {
    const double g       = r->degree_oblique;
    int          delta_j = j - r->NY / 2;
    double       sign    = (g >= 0.0 ? (double)delta_j : -(double)delta_j);
    if (!(delta_j == 0)) sign = sign / fabs(sign);
    int io           = (int)((int)(sign * g * delta_j) * sign);
    *alpha           = (g * delta_j - io) * sign;
    *beta            = 1.0 - *alpha;
    io               = (int)fmod(i + io + 4.0 * r->NX, r->NX);
    *i_oblique       = io;
    *i_oblique_plus  = (int)fmod(io + sign + 4.0 * r->NX, r->NX);
}

/* U_oblique2u (fft_wrapper.h:480-505): oblique grid and contravariant
 * components -> Cartesian grid and components, plus the mean shear flow */
static void u_oblique2u(kr_run *r, double **uu, int add_mean_flow) {
    double **w = r->w3;
    for (int d = 0; d < 3; d++) memcpy(w[d], uu[d], sizeof(double) * r->M);
    const double g = r->degree_oblique, s = r->shear_rate_eff;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++) {
            int    i_plus, i_oblique;
            double alpha, beta;
            obl2orth(r, j, i, &i_plus, &i_oblique, &alpha, &beta);
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k), io = IM(r, i_oblique, j, k), ip = IM(r, i_plus, j, k);
                for (int d = 0; d < 3; d++) uu[d][io] = beta * w[d][im] + alpha * w[d][ip];
                uu[0][io] += (g * uu[1][io]);
                if (add_mean_flow) uu[0][io] += s * (j - r->NY / 2.0);
            }
        }
}

/* U2u_oblique (fft_wrapper.h:448-478): Cartesian -> oblique (the mean
 * shear flow is not removed) */
static void u2u_oblique(kr_run *r, double **uu) {
    double **w = r->w3;
    for (int d = 0; d < 3; d++) memcpy(w[d], uu[d], sizeof(double) * r->M);
    const double g = r->degree_oblique;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++) {
            int    i_oblique, i_oblique_plus;
            double alpha, beta;
            orth2obl(r, j, i, &i_oblique, &i_oblique_plus, &alpha, &beta);
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k), io = IM(r, i_oblique, j, k), ip = IM(r, i_oblique_plus, j, k);
                for (int d = 0; d < 3; d++) uu[d][im] = beta * w[d][io] + alpha * w[d][ip];
                uu[0][im] -= (g * uu[1][im]);
            }
        }
}

/* adj (fdm.h:52) */
static int adj(int x, int i, int n) { return ((i + x) < 0) ? i + x + n : (((i + x) > n - 1) ? i + x - n : i + x); }

/* Calc_shear_rate_eff (sp_3d_ns.h:189-203) with Calc_local_gradient_y_OBL
 * (sp_3d_ns.h:166-187), summed serially in (i, j, k) order */
static void calc_shear_rate_eff(kr_run *r) {
    const double  g = r->degree_oblique, inv_2dx = 1. / (2. * r->DX);
    const double *f = r->u[0];
    double        s = 0.0;
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                int    ip1 = adj(1, i, r->NX), jp1 = adj(1, j, r->NY), im1 = adj(-1, i, r->NX), jm1 = adj(-1, j, r->NY);
                double fp1 = 0.5 * ((1 - g) * f[IM(r, ip1, jp1, k)] + (1 + g) * f[IM(r, im1, jp1, k)]);
                double fm1 = 0.5 * ((1 - g) * f[IM(r, im1, jm1, k)] + (1 + g) * f[IM(r, ip1, jm1, k)]);
                s += (fp1 - fm1) * inv_2dx;
            }
    s *= r->Ivolume * r->DX3;
    r->shear_rate_eff += s;
}

/* Solenoidal_uk_OBL (operate_omega.cxx:197-227) */
static void solenoidal_uk_obl(kr_run *r, double **u) {
    const double g = r->degree_oblique;
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ_; k++) {
                size_t im = IM(r, i, j, k);
                double kx = r->KX[im] * r->WAVE[0], ky = r->KY[im] * r->WAVE[1], kz = r->KZ[im] * r->WAVE[2];
                double kxc = (1. + g * g) * r->KX[im] * r->WAVE[0] - g * r->KY[im] * r->WAVE[1];
                double kyc = -g * r->KX[im] * r->WAVE[0] + r->KY[im] * r->WAVE[1];
                double dmy = r->IK2[im] * (u[0][im] * kx + u[1][im] * ky + u[2][im] * kz);
                u[0][im] -= dmy * kxc;
                u[1][im] -= dmy * kyc;
                u[2][im] -= dmy * kz;
            }
}

/* U_k2zeta_k_OBL (fft_wrapper.cxx:526-589) */
static void u_k2zeta_k_obl(kr_run *r, double **u, double **zeta, double *uk_dc) {
    const double g = r->degree_oblique;
    uk_dc[0]       = u[0][0];
    uk_dc[1]       = u[1][0];
    uk_dc[2]       = u[2][0];
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int kk = 0; kk < r->HNZ_; kk++) {
                size_t im    = IM(r, i, j, 2 * kk);
                double ks[3] = {r->KX[im] * r->WAVE[0], r->KY[im] * r->WAVE[1], r->KZ[im] * r->WAVE[2]};
                double ur[3] = {u[0][im], u[1][im], u[2][im]}, ui[3] = {u[0][im + 1], u[1][im + 1], u[2][im + 1]};
                contra2co_single(g, ur);
                contra2co_single(g, ui);
                double ore[3], oim[3];
                ore[0] = -(ks[1] * ui[2] - ks[2] * ui[1]);
                oim[0] = (ks[1] * ur[2] - ks[2] * ur[1]);
                ore[1] = -(ks[2] * ui[0] - ks[0] * ui[2]);
                oim[1] = (ks[2] * ur[0] - ks[0] * ur[2]);
                ore[2] = -(ks[0] * ui[1] - ks[1] * ui[0]);
                oim[2] = (ks[0] * ur[1] - ks[1] * ur[0]);
                int a0, a1;
                if (r->KX[im] != 0) {
                    a0 = 1;
                    a1 = 2;
                } else if (r->KZ[im] != 0) {
                    a0 = 1;
                    a1 = 0;
                } else {
                    a0 = 0;
                    a1 = 2;
                }
                zeta[0][im]     = ore[a0];
                zeta[0][im + 1] = oim[a0];
                zeta[1][im]     = ore[a1];
                zeta[1][im + 1] = oim[a1];
            }
}

/* PBC_OBL (periodic_boundary.h:26-44): returns the Y crossing (+1 up, -1 down) */
static int pbc_obl(const kr_run *r, double *x, double *delta_vx) {
    double signY = x[1];
    x[1]         = fmod(x[1] + r->L[1], r->L[1]);
    signY -= x[1];
    int sign = (int)signY;
    if (!(sign == 0)) sign = sign / abs(sign);
    x[0] -= (double)sign * r->degree_oblique * r->L[1];
    x[0]      = fmod(x[0] + r->L[0], r->L[0]);
    x[2]      = fmod(x[2] + r->L[2], r->L[2]);
    *delta_vx = -(double)sign * r->shear_rate_eff * r->L[1];
    return sign;
}

/* Distance0_OBL (interaction.h:50-72) */
static void distance0_obl(const kr_run *r, const double *x1, const double *x2, double *r12, double *x12) {
    double dmy   = 0.0;
    double signY = x2[1] - x1[1];
    x12[1]       = x2[1] - x1[1];
    x12[1] -= (double)Nint(x12[1] / r->L[1]) * r->L[1];
    signY -= x12[1];
    int sign = (int)signY;
    if (!(sign == 0)) sign = sign / abs(sign);
    dmy += SQ(x12[1]);
    x12[0] = x2[0] - (x1[0] + (double)sign * r->degree_oblique * r->L[1]);
    x12[0] -= (double)Nint(x12[0] / r->L[0]) * r->L[0];
    dmy += SQ(x12[0]);
    x12[2] = x2[2] - x1[2];
    x12[2] -= (double)Nint(x12[2] / r->L[2]) * r->L[2];
    dmy += SQ(x12[2]);
    *r12 = sqrt(dmy);
}
static double distance_obl(const kr_run *r, const double *x1, const double *x2) {
    double d, v[3];
    distance0_obl(r, x1, x2, &d, v);
    return d;
}

/* Relative_coord_check_stepover_Y (make_phi.h:229-268): the Sekibun point
 * on the oblique grid; returns the Y crossing of the point */
static int relative_coord_obl(const kr_run *r, const int *cell, const int *x_int, const double *residue, int *r_mesh,
                              double *rv) {
    const int N[3]  = {r->NX, r->NY, r->NZ};
    double    signY = x_int[1] + cell[1];
    r_mesh[1]       = (x_int[1] + cell[1] + N[1]) % N[1];
    signY -= r_mesh[1];
    int sign = (int)signY;
    if (!(sign == 0)) sign = sign / abs(sign);
    double xo = fmod(x_int[0] - sign * r->degree_oblique * N[1] + residue[0] / r->DX + cell[0] + 4 * N[0], N[0]);
    r_mesh[0] = (int)xo;
    double xr = xo - r_mesh[0];
    r_mesh[2] = (x_int[2] + cell[2] + N[2]) % N[2];
    rv[0]     = (double)cell[0] * r->DX - xr * r->DX;
    rv[1]     = (double)cell[1] * r->DX - residue[1];
    rv[2]     = (double)cell[2] * r->DX - residue[2];
    return sign;
}

/* Make_phi_particle_sum_OBL (make_phi.cxx:262-311, 498-502) */
static void make_phi_particle_sum_obl(kr_run *r, double *phi, double *phi_sum) {
    for (int n = 0; n < r->np; n++) {
        double xp[3] = {r->p[n].x[0], r->p[n].x[1], r->p[n].x[2]};
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3];
            relative_coord_obl(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            phi_sum[IM(r, r_mesh[0], r_mesh[1], r_mesh[2])] += Phi(r, distance_obl(r, x, xp));
        }
    }
#pragma omp parallel for
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++)
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                phi[im]   = fmin(phi_sum[im], 1.0);
            }
}

/* Make_u_particle_sum_OBL (make_phi.cxx:356-402, 510-514): a point across
 * the Y boundary sees the particle's velocity shifted by the shear jump */
static void make_u_particle_sum_obl(kr_run *r, double **up, const double *phi_sum) {
    const double jump = r->shear_rate_eff * r->L[1];
    for (int n = 0; n < r->np; n++) {
        const Particle *pn = &r->p[n];
        double          xp[3], vp[3], om[3];
        for (int d = 0; d < 3; d++) {
            xp[d] = pn->x[d];
            vp[d] = pn->v[d];
            om[d] = pn->omega[d];
        }
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], x[3], v_rot[3];
            int    sign = relative_coord_obl(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            for (int d = 0; d < 3; d++) x[d] = r_mesh[d] * r->DX;
            size_t im      = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            double dmy_phi = Phi(r, distance_obl(r, x, xp)) / fmax(phi_sum[im], 1.0);
            angular2v(r, om, rv, v_rot);
            up[0][im] += ((vp[0] - sign * jump + v_rot[0]) * dmy_phi);
            up[1][im] += (vp[1] + v_rot[1]) * dmy_phi;
            up[2][im] += (vp[2] + v_rot[2]) * dmy_phi;
        }
    }
}

/* Calc_f_hydro_correct_precision_OBL (md_force.cxx:508-745), non-rigid;
 * the distance is |r| of the stepover-corrected relative coordinate. The
 * stress fields Hydro_force / Hydro_force_new are summed as they are
 * accumulated (the reference sums them over the grid afterwards,
 * Calc_hydro_stress, fluid_solver.h:253-279): per particle in Sekibun
 * order, particles in index order. */
static void calc_f_hydro_obl(kr_run *r, const double *phi_sum, double **u) {
    const double dmy = (-r->DX3 * r->RHO) / r->jikan.dt_fluid, jump = r->shear_rate_eff * r->L[1];
    double      *hf_old = r->le_hf, *hf_new = r->le_hf + r->np;
    double      *vol = r->le_hf + 2 * r->np, *itr = r->le_hf + 3 * r->np, *f0 = r->le_hf + 4 * r->np;
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *pn = &r->p[n];
        double    xp[3], vp[3], om[3], force[3] = {0, 0, 0}, torque[3] = {0, 0, 0};
        double    rhop = r->spec[pn->spec].mass_ratio * r->RHO; /* RHO_particle, input.cxx:430 */
        for (int d = 0; d < 3; d++) {
            xp[d] = pn->x[d];
            vp[d] = pn->v[d];
            om[d] = pn->omega[d];
        }
        int    x_int[3];
        double residue[3];
        particle_cell(r, xp, x_int, residue);
        double a = 0.0, vn = 0.0, it = 0.0;
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3], v_rot[3], fp[3];
            int    sign = relative_coord_obl(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            double dmyR = sqrt(SQ(rv[0]) + SQ(rv[1]) + SQ(rv[2]));
            size_t im   = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            double dphi = Phi(r, dmyR) / fmax(phi_sum[im], 1.0);
            angular2v(r, om, rv, v_rot);
            fp[0] = ((vp[0] - sign * jump + v_rot[0]) - u[0][im]) * dphi;
            fp[1] = ((vp[1] + v_rot[1]) - u[1][im]) * dphi;
            fp[2] = ((vp[2] + v_rot[2]) - u[2][im]) * dphi;
            for (int d = 0; d < 3; d++) force[d] += fp[d];
            torque[0] += (rv[1] * fp[2] - rv[2] * fp[1]);
            torque[1] += (rv[2] * fp[0] - rv[0] * fp[2]);
            torque[2] += (rv[0] * fp[1] - rv[1] * fp[0]);
            double ry = (r_mesh[1] + sign * r->L[1]);
            a += fp[0] * ry * rhop; /* Hydro_force and Hydro_force_new, md_force.cxx:620-623 */
            vn += dphi;
            it += dphi * SQ(dmyR);
        }
        it *= 2.0 / 3.0;
        for (int d = 0; d < 3; d++) pn->f_hydro[d] = (dmy * force[d]);
        if (r->P.rotation)
            for (int d = 0; d < 3; d++) pn->torque_hydro[d] = (dmy * torque[d]);
        /* md_force.cxx:655-679: the momentum and rigid-motion corrections */
        double ho = a, hn = a;
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3];
            int    sign = relative_coord_obl(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            double dmyR = sqrt(SQ(rv[0]) + SQ(rv[1]) + SQ(rv[2]));
            size_t im   = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            double dphi = Phi(r, dmyR) / fmax(phi_sum[im], 1.0);
            double ry   = (r_mesh[1] + sign * r->L[1]);
            ho -= (pn->momentum_depend_fr[0] / vn) * ry * dphi;
            double sv = -r->RHO * force[0] / vn;
            double sw = -r->RHO * (torque[1] * rv[2] - torque[2] * rv[1]) / it;
            hn += (sv + sw) * ry * dphi;
        }
        hf_old[n] = ho;
        hf_new[n] = hn;
        vol[n]    = vn;
        itr[n]    = it;
        f0[n]     = force[0];
    }
    /* md_force.cxx:681-745: the mean force, spread back by volume */
    double sum_force = 0.0, sum_volume = 0.0;
    for (int n = 0; n < r->np; n++) {
        sum_force += f0[n];
        sum_volume += vol[n];
    }
    sum_volume /= r->RHO;
    double old = 0.0, nw = 0.0;
    for (int n = 0; n < r->np; n++) {
        const Particle *pn = &r->p[n];
        double          xp[3] = {pn->x[0], pn->x[1], pn->x[2]}, residue[3], b = 0.0;
        int             x_int[3];
        particle_cell(r, xp, x_int, residue);
        for (int mesh = 0; mesh < r->NP_domain; mesh++) {
            int    r_mesh[3];
            double rv[3];
            int    sign = relative_coord_obl(r, r->sek + 3 * mesh, x_int, residue, r_mesh, rv);
            double dmyR = sqrt(SQ(rv[0]) + SQ(rv[1]) + SQ(rv[2]));
            size_t im   = IM(r, r_mesh[0], r_mesh[1], r_mesh[2]);
            double dphi = Phi(r, dmyR) / fmax(phi_sum[im], 1.0);
            b -= sum_force * (r_mesh[1] + sign * r->L[1]) * dphi / sum_volume;
        }
        old += hf_old[n] + b;
        nw += hf_new[n];
    }
    /* Calc_hydro_stress (fluid_solver.h:253-279) */
    const double k = (r->Ivolume * r->DX3) / r->jikan.dt_fluid;
    r->le_stress_old = -old * k;
    r->le_stress_new = -nw * k;
}

/* Calc_Reynolds_shear_stress (fluid_solver.h:165-183) on the Cartesian u */
static void calc_reynolds_shear_stress(kr_run *r, double **u) {
    double s = 0.0;
    for (int i = 0; i < r->NX; i++)
        for (int j = 0; j < r->NY; j++) {
            double dy = (double)(j - r->NY / 2) * r->DX;
            for (int k = 0; k < r->NZ; k++) {
                size_t im = IM(r, i, j, k);
                s += (u[0][im] - r->shear_rate_eff * dy) * (u[1][im]);
            }
        }
    r->le_reynolds = -(r->Ivolume * r->DX3) * r->RHO * s;
}

/* Calc_f_Lennard_Jones_shear_cap_primitive (md_force.cxx:198-305) with
 * Distance0_OBL and cap = DBL_MAX (md_force.h:82-84): pairs m > n, each
 * particle's own sum added after its row */
static void calc_f_lennard_jones_obl(kr_run *r) {
    const double pair_cutoff = r->A_R_cutoff * r->LJ_dia;
    double       ss0         = 0.0;
    for (int n = 0; n < r->np; n++) {
        Particle *pn     = &r->p[n];
        double    sum[3] = {0, 0, 0};
        for (int m = n + 1; m < r->np; m++) {
            Particle *pm = &r->p[m];
            if (is_obstacle(r, pn) && is_obstacle(r, pm)) continue;
            double rij, rv[3];
            distance0_obl(r, pn->x, pm->x, &rij, rv);
            if (rij < pair_cutoff) {
                double dmy_r = fmin(DBL_MAX / rij, lennard_jones_f(r, rij));
                for (int d = 0; d < 3; d++) {
                    double fn = dmy_r * (-rv[d]);
                    pm->fr[d] -= fn;
                    sum[d] += fn;
                }
                ss0 += (dmy_r * (-rv[0])) * rv[1]; /* the pair's shear stress, md_force.cxx:246 */
            }
        }
        for (int d = 0; d < 3; d++) pn->fr[d] += sum[d];
    }
    r->le_dev_lj += ss0;
}

/* MD_solver_position_Euler_OBL / _AB2_OBL (particle_solver.cxx:283-334) */
static void position_update_obl(kr_run *r, int ab2) {
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p = &r->p[n];
        if (is_obstacle(r, p)) continue;
        for (int d = 0; d < 3; d++) {
            p->x_previous[d] = p->x[d];
            double delta_x   = ab2 ? r->jikan.hdt_md * (3.0 * p->v[d] - p->v_old[d]) : r->jikan.dt_md * p->v[d];
            p->x_nopbc[d] += delta_x;
            p->x[d] += delta_x;
        }
        double delta_vx;
        pbc_obl(r, p->x, &delta_vx);
        p->v[0] += delta_vx;
        p->v_old[0] += delta_vx;
        if (r->P.rotation) orientation_update(r, p, ab2);
    }
}

/* MD_solver_velocity_Euler_OBL / _AB2_hydro_OBL (particle_solver.cxx:336-432)
 * with Force_OBL (particle_solver.h:113-126) */
static void velocity_update_obl(kr_run *r, int ab2) {
    r->le_dev_lj = 0.0; /* Force_OBL, particle_solver.h:113-126 */
    if (r->P.lj_truncate >= 0) calc_f_lennard_jones_obl(r);
    if (r->P.g != 0.0) add_f_gravity(r);
    r->le_dev_lj *= r->Ivolume;
#pragma omp parallel for
    for (int n = 0; n < r->np; n++) {
        Particle *p    = &r->p[n];
        int       s    = p->spec;
        double    h    = ab2 ? r->jikan.hdt_md : r->jikan.dt_md;
        double    dmy  = h * r->IMASS[s];
        double    dmyr = h * r->IMOI[s];
        if (!is_obstacle(r, p)) {
            double sf[3], st[3];
            self_propulsion(r, p, sf, st);
            for (int d = 0; d < 3; d++) {
                p->v_old[d]     = p->v[d];
                p->omega_old[d] = p->omega[d];
                if (!ab2) {
                    p->v[d] += (dmy * (p->f_hydro[d] + p->fr[d] + sf[d]));
                    p->omega[d] += (dmyr * (p->torque_hydro[d] + p->torque_r[d] + st[d]));
                } else {
                    p->v[d] += dmy * (2.0 * (p->f_hydro[d] + sf[d]) + p->fr[d] + p->fr_previous[d]);
                    p->omega[d] +=
                        dmyr * (2.0 * (p->torque_hydro[d] + st[d]) + p->torque_r[d] + p->torque_r_previous[d]);
                }
            }
        } else {
            for (int d = 0; d < 3; d++) {
                p->v_old[d]     = p->v[d];
                p->v[d]         = 0.0;
                p->omega_old[d] = p->omega[d];
                p->omega[d]     = 0.0;
            }
        }
        for (int d = 0; d < 3; d++)
            p->momentum_depend_fr[d] = ab2 ? h * (p->fr[d] + p->fr_previous[d]) : h * p->fr[d];
    }
    reset_forces(r);
}

/* Time_evolution_hydro_OBL (sp_3d_ns.cxx:328-433) */
static void time_evolution_hydro_obl(kr_run *r) {
    ns_solver_obl(r);
    if (r->fix_cell_any)
        for (int d = 0; d < 3; d++)
            if (r->P.fix_cell[d]) r->uk_dc[d] = 0.0;
    zeta_k2u_k_obl(r, r->zeta, r->uk_dc, r->u);
    for (int d = 0; d < 3; d++) A_k2a(r, r->u[d]);
    if (!r->le_ac)
        r->shear_rate_eff = r->shear_rate;
    else
        r->shear_rate_eff = r->shear_rate * cos(2. * KR_PI * r->shear_freq * r->jikan.dt_fluid * r->jikan.ts);
    r->degree_oblique += r->shear_rate_eff * r->jikan.dt_fluid;
    if (r->degree_oblique >= r->XYaspect || r->degree_oblique <= -r->XYaspect) {
        int flg = r->degree_oblique >= r->XYaspect ? 0 : 1;
        reset_u_obl(r, r->ucp, r->u, flg);
        for (int d = 0; d < 3; d++) {
            double *t = r->u[d];
            r->u[d]   = r->ucp[d];
            r->ucp[d] = t;
        }
        r->degree_oblique += flg == 0 ? -r->XYaspect : r->XYaspect;
    }
    update_k2_obl(r);
    for (int d = 0; d < 3; d++) memcpy(r->ucp[d], r->u[d], sizeof(double) * r->M);
    u_oblique2u(r, r->ucp, 1);
    if (!r->le_ac) calc_shear_rate_eff(r);

    position_update_obl(r, r->jikan.ts != 0);
    reset_field(r, r->phi);
    reset_field(r, r->phi_sum);
    for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
    make_phi_particle_sum_obl(r, r->phi, r->phi_sum);
    calc_f_hydro_obl(r, r->phi_sum, r->ucp);
    calc_reynolds_shear_stress(r, r->ucp);
    velocity_update_obl(r, r->jikan.ts != 0);
    if (r->kBT > 0.) thermostat(r);
    for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
    make_u_particle_sum_obl(r, r->up, r->phi_sum);
    make_f_particle_nonsole(r, r->f, r->ucp, r->up, r->phi);
    u2u_oblique(r, r->f);
    add_f_particle(r, r->u, r->f);
    for (int d = 0; d < 3; d++) A2a_k(r, r->u[d]);
    solenoidal_uk_obl(r, r->u);
    u_k2zeta_k_obl(r, r->u, r->zeta, r->uk_dc);
}

/* Mean_shear_stress SHOW for Shear_Navier_Stokes_Lees_Edwards (sp_3d_ns.h:247-330),
 * after the step and its time increment, with the strain update of
 * sp_3d_ns.cxx:962-964 */
static void le_shear_record(kr_run *r) {
    r->le_strain += r->shear_rate_eff * r->jikan.dt_fluid;
    const double fluid    = r->ETA * r->shear_rate_eff;
    const double apparent = r->le_stress_new + r->le_reynolds + r->le_dev_lj + fluid + 0.0;
    const double rec[KR_LE_RECORD] = {r->jikan.time,  r->shear_rate_eff, r->degree_oblique, r->le_strain,
                                      r->le_dev_lj,   r->le_stress_old,  r->le_stress_new,  r->le_reynolds,
                                      fluid,          0.0,               apparent,          apparent / r->shear_rate_eff};
    memcpy(r->le_record, rec, sizeof(rec));
}
void kr_get_shear_record(kr_run *r, double out[KR_LE_RECORD]) { memcpy(out, r->le_record, sizeof(r->le_record)); }
void kr_get_shear_diag(kr_run *r, double out[4]) {
    out[0] = r->le_stress_old;
    out[1] = r->le_stress_new;
    out[2] = r->le_reynolds;
    out[3] = r->le_dev_lj;
}

int kr_set_shear(kr_run *r, double shear_rate, int ac, double frequency) {
    if (!r) return -1;
    if (!r->le) {
        for (int d = 0; d < 3; d++) {
            r->ucp[d] = (double *)calloc(r->M, sizeof(double));
            r->w3[d]  = (double *)calloc(r->M, sizeof(double));
        }
        r->w1    = (double *)calloc(r->M, sizeof(double));
        r->le_hf = (double *)calloc(5 * (size_t)(r->np > 0 ? r->np : 1), sizeof(double));
    }
    r->le         = 1;
    r->le_ac      = ac ? 1 : 0;
    r->shear_rate = shear_rate;
    r->shear_freq = frequency;
    r->XYaspect   = (double)r->NX / (double)r->NY; /* sp_3d_ns.cxx:772 */
    r->Ivolume    = 1. / (r->L[0] * r->L[1] * r->L[2]);
    /* input.cxx:549-566: the sheared runs' LJ diameter */
    double dummy_pow = (r->P.lj_powers == 0) ? pow(2., 1. / 6.) : (r->P.lj_powers == 1) ? pow(2., 1. / 12.)
                                                                                        : pow(2., 1. / 18.);
    r->LJ_dia = fmin((r->SIGMA + r->XI) / dummy_pow, r->SIGMA);
    update_k2_obl(r);
    return 0;
}
void kr_get_shear_state(kr_run *r, double *degree_oblique, double *shear_rate_eff) {
    *degree_oblique = r->degree_oblique;
    *shear_rate_eff = r->shear_rate_eff;
}
/* resume.cxx:566-580: the strain and the K2 it implies */
void kr_set_shear_state(kr_run *r, double degree_oblique, double shear_rate_eff) {
    r->degree_oblique = degree_oblique;
    r->shear_rate_eff = shear_rate_eff;
    update_k2_obl(r);
}

int kr_le_apply(kr_run *r, int fn) {
    if (!r->le) return -1;
    switch (fn) {
        case KR_LE_OBLIQUE2U: u_oblique2u(r, r->u, 1); break;
        case KR_LE_U2OBLIQUE: u2u_oblique(r, r->u); break;
        case KR_LE_RESET_UP:
        case KR_LE_RESET_DOWN:
            reset_u_obl(r, r->ucp, r->u, fn == KR_LE_RESET_UP ? 0 : 1);
            for (int d = 0; d < 3; d++) {
                double *t = r->u[d];
                r->u[d]   = r->ucp[d];
                r->ucp[d] = t;
            }
            break;
        case KR_LE_ZETA_K2U:
            zeta_k2u_k_obl(r, r->zeta, r->uk_dc, r->u);
            for (int d = 0; d < 3; d++) A_k2a(r, r->u[d]);
            break;
        case KR_LE_U2ZETA_K:
            for (int d = 0; d < 3; d++) A2a_k(r, r->u[d]);
            u_k2zeta_k_obl(r, r->u, r->zeta, r->uk_dc);
            break;
        case KR_LE_SOLENOIDAL_U:
            for (int d = 0; d < 3; d++) A2a_k(r, r->u[d]);
            solenoidal_uk_obl(r, r->u);
            for (int d = 0; d < 3; d++) A_k2a(r, r->u[d]);
            break;
        case KR_LE_NS_SOLVER: ns_solver_obl(r); break;
        case KR_LE_SHEAR_RATE_EFF: calc_shear_rate_eff(r); break;
        case KR_LE_POSITION_EULER: position_update_obl(r, 0); break;
        case KR_LE_MAKE_PHI:
            reset_field(r, r->phi);
            reset_field(r, r->phi_sum);
            make_phi_particle_sum_obl(r, r->phi, r->phi_sum);
            break;
        case KR_LE_MAKE_UP:
            for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
            make_u_particle_sum_obl(r, r->up, r->phi_sum);
            break;
        case KR_LE_F_HYDRO: calc_f_hydro_obl(r, r->phi_sum, r->u); break;
        case KR_LE_LJ:
            r->le_dev_lj = 0.0;
            calc_f_lennard_jones_obl(r);
            r->le_dev_lj *= r->Ivolume;
            break;
        case KR_LE_REYNOLDS: calc_reynolds_shear_stress(r, r->u); break;
        default: return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* the step and its initial block                                       */
/* ------------------------------------------------------------------ */
/* Time_evolution_hydro (sp_3d_ns.cxx:41-194), Navier_Stokes branch */
static void time_evolution_hydro(kr_run *r) {
    const int ab2 = r->jikan.ts != 0 || g_mut == KR_MUT_NO_EULER_START;
    if (g_mut == KR_MUT_FH_PRE_NS_U) zeta_k2u(r, r->zeta, r->uk_dc, r->f); /* u before the NS update */
    ns_solver_slaved_euler(r);
    if (r->fix_cell_any && g_mut != KR_MUT_NO_FIX_CELL)
        for (int d = 0; d < 3; d++)
            if (r->P.fix_cell[d]) r->uk_dc[d] = 0.0;
    truncate_two_third(r, r->zeta[0]);
    truncate_two_third(r, r->zeta[1]);
    zeta_k2u(r, r->zeta, r->uk_dc, r->u);
    if (g_mut == KR_MUT_FH_OLD_X) { /* F_H at the positions before the move */
        reset_field(r, r->phi);
        reset_field(r, r->phi_sum);
        make_phi_particle_sum(r, r->phi, r->phi_sum);
        calc_f_hydro(r, r->phi_sum, r->u);
    }
    position_update(r, ab2);
    reset_field(r, r->phi);
    reset_field(r, r->phi_sum);
    make_phi_particle_sum(r, r->phi, r->phi_sum);
    for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
    if (g_mut != KR_MUT_FH_OLD_X) calc_f_hydro(r, r->phi_sum, g_mut == KR_MUT_FH_PRE_NS_U ? r->f : r->u);
    if (!r->sw_janus_slip)
        velocity_update(r, ab2);
    else
        slip_iteration(r);
    if (r->kBT > 0.) thermostat(r);
    for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
    make_u_particle_sum(r, r->up, r->phi_sum);
    make_f_particle_nonsole(r, r->f, r->u, r->up, r->phi);
    solenoidal_u(r, r->f);
    add_f_particle(r, r->u, r->f);
    u2zeta_k(r, r->zeta, r->uk_dc, r->u);
}

void kr_step(kr_run *r, int nsteps) {
    for (int s = 0; s < nsteps; s++) {
        if (r->le)
            time_evolution_hydro_obl(r);
        else
            time_evolution_hydro(r);
        r->jikan.time += r->jikan.dt_fluid;
        r->jikan.ts++;
        if (r->le) le_shear_record(r);
    }
}

/* Init_zeta_k (init_fluid.cxx:10-25) on a fluid at rest, then the initial
 * constraint block (sp_3d_ns.cxx:809-826). */
void kr_init_fluid(kr_run *r) {
    for (int d = 0; d < 3; d++) reset_field(r, r->u[d]);
    for (int d = 0; d < 3; d++) A2a_k(r, r->u[d]);
    solenoidal_uk(r, r->u);
    u_k2zeta_k(r, r->u, r->zeta, r->uk_dc);

    reset_field(r, r->phi);
    for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
    reset_field(r, r->phi_sum);
    make_phi_particle_sum(r, r->phi, r->phi_sum);
    make_u_particle_sum(r, r->up, r->phi_sum);
    zeta_k2u(r, r->zeta, r->uk_dc, r->u);
    make_f_particle_nonsole(r, r->f, r->u, r->up, r->phi);
    solenoidal_u(r, r->f);
    add_f_particle(r, r->u, r->f);
    u2zeta_k(r, r->zeta, r->uk_dc, r->u);
    for (int d = 0; d < 3; d++) r->uk_dc[d] = 0.0;
}

int kr_apply(kr_run *r, int fn) {
    switch (fn) {
        case KH_FN_NS_SOLVER: ns_solver_slaved_euler(r); break;
        case KH_FN_ZETA_K2U: zeta_k2u(r, r->zeta, r->uk_dc, r->u); break;
        case KH_FN_U2ZETA_K: u2zeta_k(r, r->zeta, r->uk_dc, r->u); break;
        case KH_FN_POSITION_EULER: position_update(r, 0); break;
        case KH_FN_POSITION_AB2: position_update(r, 1); break;
        case KH_FN_MAKE_PHI:
            reset_field(r, r->phi);
            reset_field(r, r->phi_sum);
            make_phi_particle_sum(r, r->phi, r->phi_sum);
            break;
        case KH_FN_F_HYDRO: calc_f_hydro(r, r->phi_sum, r->u); break;
        case KH_FN_VELOCITY_EULER: velocity_update(r, 0); break;
        case KH_FN_VELOCITY_AB2: velocity_update(r, 1); break;
        case KH_FN_THERMOSTAT: thermostat(r); break;
        case KH_FN_MAKE_UP:
            for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
            make_u_particle_sum(r, r->up, r->phi_sum);
            break;
        case KH_FN_F_PARTICLE_SOLE:
            make_f_particle_nonsole(r, r->f, r->u, r->up, r->phi);
            solenoidal_u(r, r->f);
            break;
        case KH_FN_ADD_F_PARTICLE: add_f_particle(r, r->u, r->f); break;
        case KH_FN_SLIP_MOMENTUM: make_particle_momentum_factor(r, r->u); break;
        case KH_FN_SLIP_SCATTER:
            for (int d = 0; d < 3; d++) reset_field(r, r->up[d]);
            make_force_u_slip_particle(r, r->up, r->u);
            break;
        case KH_FN_SLIP_ITERATION: slip_iteration(r); break;
        default: return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* lifecycle and transfer                                               */
/* ------------------------------------------------------------------ */
kr_run *kr_create(const kh_params *P, int np) {
    if (!P || P->nx < 2 || P->ny < 2 || P->nz < 2 || (P->nz & 1) || P->nspecies < 1 || !P->species || np < 0)
        return NULL;
    kr_run *r = (kr_run *)calloc(1, sizeof(kr_run));
    r->P      = *P;
    r->nspec  = P->nspecies;
    r->spec   = (kh_species *)malloc(sizeof(kh_species) * (size_t)r->nspec);
    memcpy(r->spec, P->species, sizeof(kh_species) * (size_t)r->nspec);
    r->P.species = r->spec;
    /* Set_global_parameters (input.cxx:424-600) */
    r->NX     = P->nx;
    r->NY     = P->ny;
    r->NZ     = P->nz;
    r->HNX    = P->nx / 2;
    r->HNY    = P->ny / 2;
    r->HNZ_   = P->nz / 2 + 1;
    r->NZ_    = 2 * r->HNZ_;
    r->TRN_X  = (P->nx + 2) / 3;
    r->TRN_Y  = (P->ny + 2) / 3;
    r->TRN_Z  = (P->nz + 2) / 3;
    r->DX     = P->dx;
    r->RHO    = P->rho;
    r->ETA    = P->eta;
    r->IRHO   = 1. / P->rho;
    r->NU     = P->eta * r->IRHO;
    r->DX3    = P->dx * P->dx * P->dx;
    r->kBT    = P->kbt;
    r->RADIUS = P->a * P->dx;
    r->XI     = P->a_xi * P->dx;
    r->SIGMA  = 2.0 * r->RADIUS;
    r->HXI    = r->XI * 0.5;
    int N[3]  = {P->nx, P->ny, P->nz};
    for (int d = 0; d < 3; d++) {
        r->L[d]    = N[d] * P->dx;
        r->WAVE[d] = 2. * KR_PI / r->L[d];
    }
    r->MASS           = (double *)malloc(sizeof(double) * (size_t)r->nspec);
    r->IMASS          = (double *)malloc(sizeof(double) * (size_t)r->nspec);
    r->MOI            = (double *)malloc(sizeof(double) * (size_t)r->nspec);
    r->IMOI           = (double *)malloc(sizeof(double) * (size_t)r->nspec);
    r->slip_mode_half = (double *)malloc(sizeof(double) * (size_t)r->nspec);
    for (int i = 0; i < r->nspec; i++) {
        double rho_p         = r->spec[i].mass_ratio * r->RHO;
        r->MASS[i]           = rho_p * 4. / 3. * KR_PI * (r->RADIUS * r->RADIUS * r->RADIUS);
        r->IMASS[i]          = 1. / r->MASS[i];
        r->MOI[i]            = 2. / 5. * r->MASS[i] * SQ(r->RADIUS);
        r->IMOI[i]           = 1. / r->MOI[i];
        r->slip_mode_half[i] = r->spec[i].slip_mode / 2.0;
        if (r->spec[i].propulsion == KH_JP_MOTOR) r->sw_janus_motor = 1;
        if (r->spec[i].propulsion == KH_JP_SLIP) r->sw_janus_slip = 1;
    }
    r->fix_cell_any = P->fix_cell[0] | P->fix_cell[1] | P->fix_cell[2];
    /* LJ: input.cxx:1798-1826 and 559-569 (non-shear: LJ_dia = SIGMA) */
    if (P->lj_truncate > 0)
        r->A_R_cutoff = (P->lj_powers == 0) ? pow(2., 1. / 6.) : (P->lj_powers == 1) ? pow(2., 1. / 12.) : pow(2., 1. / 18.);
    else if (P->lj_truncate == 0) {
        int nmin      = P->nx < P->ny ? P->nx : P->ny;
        nmin          = nmin < P->nz ? nmin : P->nz;
        r->A_R_cutoff = fmin(nmin * P->dx * .5 / r->SIGMA, 2.5);
    } else
        r->A_R_cutoff = 0.;
    r->LJ_dia           = r->SIGMA;
    r->Gravity_on_fluid = P->g * r->RHO * 4. / 3. * KR_PI * SQ(r->RADIUS) * r->RADIUS; /* md_force.cxx:307 */

    r->M = (size_t)r->NX * r->NY * r->NZ_;
    double **all[] = {&r->zeta[0], &r->zeta[1], &r->u[0], &r->u[1], &r->u[2], &r->up[0], &r->up[1], &r->up[2],
                      &r->f[0],    &r->f[1],    &r->phi,  &r->phi_sum, &r->f_ns0[0], &r->f_ns0[1], &r->f[2],
                      &r->K2,      &r->IK2};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); a++) *all[a] = (double *)calloc(r->M, sizeof(double));
    r->KX = (int *)calloc(r->M, sizeof(int));
    r->KY = (int *)calloc(r->M, sizeof(int));
    r->KZ = (int *)calloc(r->M, sizeof(int));
    init_K(r);
    if (kr_fft3_init(&r->fft, r->NX, r->NY, r->NZ)) {
        kr_destroy(r);
        return NULL;
    }
    particle_domain(r);
    r->np = np;
    r->p  = (Particle *)calloc((size_t)(np > 0 ? np : 1), sizeof(Particle));
    for (int n = 0; n < np; n++) r->p[n].q.s = 1.0;
    CTime t0 = {0, 0.0, P->dt, P->dt * 0.5, P->dt, P->dt * 0.5}; /* sp_3d_ns.cxx:782 */
    r->jikan = t0;
    return r;
}

void kr_destroy(kr_run *r) {
    if (!r) return;
    double *all[] = {r->zeta[0], r->zeta[1], r->u[0], r->u[1], r->u[2], r->up[0], r->up[1], r->up[2], r->f[0],
                     r->f[1],    r->f[2],    r->phi,  r->phi_sum, r->f_ns0[0], r->f_ns0[1], r->K2, r->IK2};
    for (size_t a = 0; a < sizeof(all) / sizeof(all[0]); a++) free(all[a]);
    for (int d = 0; d < 3; d++) {
        free(r->ucp[d]);
        free(r->w3[d]);
    }
    free(r->w1);
    free(r->le_hf);
    free(r->KX);
    free(r->KY);
    free(r->KZ);
    free(r->sek);
    free(r->p);
    free(r->spec);
    free(r->MASS);
    free(r->IMASS);
    free(r->MOI);
    free(r->IMOI);
    free(r->slip_mode_half);
    free(r->bcum);
    free(r->bid);
    free(r->body);
    free(r->GR);
    free(r->GRb);
    kr_fft3_free(&r->fft);
    free(r);
}

int kr_set_particles(kr_run *r, const Particle *p, int n) {
    if (n != r->np) return -1;
    for (int i = 0; i < n; i++)
        if (p[i].spec < 0 || p[i].spec >= r->nspec) return -1;
    memcpy(r->p, p, sizeof(Particle) * (size_t)n);
    return 0;
}
int kr_get_particles(kr_run *r, Particle *p, int n) {
    if (n != r->np) return -1;
    memcpy(p, r->p, sizeof(Particle) * (size_t)n);
    return 0;
}
void kr_set_fluid(kr_run *r, const double *z0, const double *z1, const double uk_dc[3]) {
    memcpy(r->zeta[0], z0, sizeof(double) * r->M);
    memcpy(r->zeta[1], z1, sizeof(double) * r->M);
    memcpy(r->uk_dc, uk_dc, sizeof(double) * 3);
}
void kr_get_fluid(kr_run *r, double *z0, double *z1, double uk_dc[3]) {
    memcpy(z0, r->zeta[0], sizeof(double) * r->M);
    memcpy(z1, r->zeta[1], sizeof(double) * r->M);
    memcpy(uk_dc, r->uk_dc, sizeof(double) * 3);
}
void kr_set_time(kr_run *r, const CTime *t) { r->jikan = *t; }
void kr_get_time(kr_run *r, CTime *t) { *t = r->jikan; }

static double *field_ptr(kr_run *r, int field) {
    switch (field) {
        case KH_FIELD_U0: return r->u[0];
        case KH_FIELD_U1: return r->u[1];
        case KH_FIELD_U2: return r->u[2];
        case KH_FIELD_UP0: return r->up[0];
        case KH_FIELD_UP1: return r->up[1];
        case KH_FIELD_UP2: return r->up[2];
        case KH_FIELD_PHI: return r->phi;
        case KH_FIELD_PHI_SUM: return r->phi_sum;
        case KH_FIELD_F0: return r->f[0];
        case KH_FIELD_F1: return r->f[1];
        case KH_FIELD_F2: return r->f[2];
        default: return NULL;
    }
}
int kr_get_field(kr_run *r, int field, double *out) {
    double *a = field_ptr(r, field);
    if (!a) return -1;
    memcpy(out, a, sizeof(double) * r->M);
    return 0;
}
int kr_set_field(kr_run *r, int field, const double *in) {
    double *a = field_ptr(r, field);
    if (!a) return -1;
    memcpy(a, in, sizeof(double) * r->M);
    return 0;
}

int kr_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void kr_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
