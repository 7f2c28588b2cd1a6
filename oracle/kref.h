/*
 * kref.h -- the oracle: a CPU FP64 restatement of KAPSEL's SPM step.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (kapsel3_amd/) links,
 * loads or calls this code; only tests/, the smoke check and the CPU-baseline
 * leg of bench.py use it: to check results and as the timed CPU baseline.
 *
 * It works on the reference's own data layout: every field is NX*NY*NZ_
 * doubles with NZ_ = 2*(NZ/2+1); real-space values live in k < NZ, spectral
 * values are (re,im) pairs at [i][j][2kz]. Each function cites the reference
 * file:line it restates. OpenMP parallelises the loops the reference
 * parallelises, plus the FFT rows.
 */
#ifndef KREF_H
#define KREF_H

#include <stdint.h>

#include "../include/kapsel_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kr_run kr_run;

kr_run *kr_create(const kh_params *params, int nparticles);
void    kr_destroy(kr_run *r);

int  kr_set_particles(kr_run *r, const Particle *p, int n);
int  kr_get_particles(kr_run *r, Particle *p, int n);
void kr_set_fluid(kr_run *r, const double *zeta0, const double *zeta1, const double uk_dc[3]);
void kr_get_fluid(kr_run *r, double *zeta0, double *zeta1, double uk_dc[3]);
void kr_set_time(kr_run *r, const CTime *t);
void kr_get_time(kr_run *r, CTime *t);
int  kr_get_field(kr_run *r, int field, double *out);
int  kr_set_field(kr_run *r, int field, const double *in);

void kr_init_fluid(kr_run *r);

/* Lees-Edwards shear (SW_EQ == Shear_Navier_Stokes_Lees_Edwards): after
 * kr_set_shear, kr_step runs Time_evolution_hydro_OBL. DC shear (ac = 0)
 * at shear_rate, or AC shear_rate*cos(2 pi frequency t). The strain
 * degree_oblique and the effective rate shear_rate_eff are run state. */
int  kr_set_shear(kr_run *r, double shear_rate, int ac, double frequency);
void kr_get_shear_state(kr_run *r, double *degree_oblique, double *shear_rate_eff);
void kr_set_shear_state(kr_run *r, double degree_oblique, double shear_rate_eff);
/* single functions of the sheared step, on the run's state */
enum {
    KR_LE_OBLIQUE2U = 0, /* U_oblique2u(u), mean flow added     fft_wrapper.h:480-505 */
    KR_LE_U2OBLIQUE,     /* U2u_oblique(u)                      fft_wrapper.h:448-478 */
    KR_LE_RESET_UP,      /* Reset_U_OBL(flag 0) + Swap_mem      fluid_solver.h:312-340 */
    KR_LE_RESET_DOWN,    /* Reset_U_OBL(flag 1) + Swap_mem */
    KR_LE_ZETA_K2U,      /* Zeta_k2u_k_OBL + U_k2u              fft_wrapper.cxx:690-768 */
    KR_LE_U2ZETA_K,      /* U2u_k + U_k2zeta_k_OBL              fft_wrapper.cxx:526-589 */
    KR_LE_SOLENOIDAL_U,  /* U2u_k + Solenoidal_uk_OBL + U_k2u   operate_omega.cxx:197-227 */
    KR_LE_NS_SOLVER,     /* NS_solver_slavedEuler_Shear_OBL     fluid_solver.cxx:411-437 */
    KR_LE_SHEAR_RATE_EFF,/* Calc_shear_rate_eff                 sp_3d_ns.h:189-203 */
    KR_LE_POSITION_EULER,/* MD_solver_position_Euler_OBL       particle_solver.cxx:283-307 */
    KR_LE_MAKE_PHI,      /* Reset_phi x2 + Make_phi_particle_sum_OBL  make_phi.cxx:262-311 */
    KR_LE_MAKE_UP,       /* Reset_u + Make_u_particle_sum_OBL   make_phi.cxx:356-402 */
    KR_LE_F_HYDRO,       /* Calc_f_hydro_correct_precision_OBL on u  md_force.cxx:508-645 */
    KR_LE_LJ,            /* Calc_f_Lennard_Jones_OBL (adds to fr)     md_force.cxx:198-305 */
    KR_LE_REYNOLDS       /* Calc_Reynolds_shear_stress on u (Cartesian) fluid_solver.h:165-183 */
};
int kr_le_apply(kr_run *r, int fn);
/* The record Mean_shear_stress prints after each sheared step
 * (sp_3d_ns.h:247-330): time, shear_rate, degree_oblique, strain,
 * lj_dev_stress, shear_stress_old, shear_stress_new, reynolds_stress,
 * fluid_stress, interfacial_stress (0), apparent_stress, viscosity. */
#define KR_LE_RECORD 12
void kr_get_shear_record(kr_run *r, double out[KR_LE_RECORD]);
/* the last stress sums: Hydro_force and Hydro_force_new stresses, the
 * Reynolds stress, the LJ stress */
void kr_get_shear_diag(kr_run *r, double out[4]);
void kr_step(kr_run *r, int nsteps);
int  kr_apply(kr_run *r, int fn);

/* A2a_k / A_k2a semantics on a reference-layout array of the run's size. */
void kr_A2a_k(kr_run *r, double *a);
void kr_A_k2a(kr_run *r, double *a);
/* Standalone transforms of any supported size (own temporary plans). */
int kr_fft_forward(int nx, int ny, int nz, double *a);
int kr_fft_inverse(int nx, int ny, int nz, double *a);

/* Smooth profile phi(r) (profile.h:30-50) and the Sekibun cell. */
double kr_phi(double r, double radius, double dx, double xi);
/* Squirmer_coord (make_phi.h:282-376) for quaternion (s, v0, v1, v2) and
 * Janus axis KH_JAX_*: unit vectors and polar/azimuthal angles of x */
void kr_squirmer_coord(const double q4[4], int axis, const double x[3], double nr[3], double nth[3], double nph[3],
                       double *theta, double *phi);
int    kr_sekibun_count(const kr_run *r);
int    kr_sekibun_cell(const kr_run *r, int *out3);

/* Philox4x32-10 block and the thermal-noise normals of one particle. */
void kr_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void kr_noise_normals(uint64_t seed, uint32_t stream, int particle, int ts, double out6[6]);

/* Rigid composite bodies (SW_PT == rigid): nb bodies, body k made of the
 * beads cum[k] .. cum[k+1]-1 (Rigid_Particle_Cumul) of the particles already
 * set, whose x_nopbc hold unwrapped coordinates of each body; free_flags[k]
 * 1 "free", 0 "fix" (Rigid_motion); v0, w0 the bodies' initial velocities
 * (nb x 3, space frame; NULL: at rest). Runs the initial block of
 * init_particle.cxx:455-487. -1 for kT > 0, slip species, sheared runs or
 * ROTATION off. */
int  kr_set_rigid(kr_run *r, int nb, const int *cum, const int *free_flags, const double *v0, const double *w0);
/* nb x KH_RIGID_WORDS doubles (the layout of kh_get_rigid) */
void kr_get_rigid(kr_run *r, double *out);

/* Planted mistakes, for the tests that show a pin catches one (CPU only).
 * kr_set_mutation(m) makes the oracle commit mistake m in its step from the
 * next call on (process-wide); KR_MUT_NONE restores it. Nothing but those
 * tests sets it. */
enum {
    KR_MUT_NONE = 0,
    KR_MUT_FH_PRE_NS_U,      /* F_H from the velocity before the NS update, not u* */
    KR_MUT_OMEGA_R_SIGN,     /* rigid velocity v - omega x r (Angular2v sign)     */
    KR_MUT_FH_NO_MAX,        /* F_H weight phi instead of phi / max(phi_sum, 1)   */
    KR_MUT_UP_NO_MAX,        /* u_p weight phi instead of phi / max(phi_sum, 1)   */
    KR_MUT_NO_FIX_CELL,      /* FIX_CELL zeroing of uk_dc skipped                 */
    KR_MUT_ADV_SIGN,         /* one advection product (u2 u3) with its sign flipped */
    KR_MUT_AB2_FH_WEIGHT,    /* AB2 velocity update with 1 F_H instead of 2 F_H   */
    KR_MUT_NO_EULER_START,   /* AB2 on the first step too                         */
    KR_MUT_FH_OLD_X,         /* F_H at the positions before the move              */
    KR_MUT_FH_SIGN,          /* F_H prefactor +rho dx^3/dt                         */
    KR_MUT_UKDC_MEAN,        /* uk_dc stored as the mean, not the sum, of u       */
    KR_MUT_SLIP_RELAX,       /* slip relaxation 0.5 instead of 0.7                */
    KR_MUT_SLIP_B2_FULL,     /* B2/B1 not halved                                  */
    KR_MUT_SLIP_FORCE_SIGN,  /* slip force with its sign flipped                  */
    KR_MUT_SLIP_FINAL_USTAR, /* after the slip loop u stays u* (no Swap_mem)      */
    KR_MUT_SLIP_ROTLET,      /* rotlet C1 doubled                                 */
    KR_MUT_SLIP_NO_RESET,    /* no velocity reset between slip iterations         */
    KR_MUT_COUNT
};
void kr_set_mutation(int m);

int kr_threads(void);
void kr_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
