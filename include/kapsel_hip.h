/*
 * kapsel_hip.h -- C ABI of the MI355X smoothed-profile-method (SPM) step.
 *
 * One kh_run holds one simulation's state resident in HBM: the reduced
 * vorticity zeta (2 spectral fields), the k=0 velocity uk_dc, the particle
 * state, and the work fields of the step. The host hands over particles and
 * fluid once, calls kh_step() for as many steps as it wants, and downloads
 * state only when it needs it (output, restart).
 *
 * Every call returns KH_OK (0) or a negative kh_status; no exception crosses
 * this boundary. kh_last_error() gives a readable message for the last
 * failure on the calling thread. Plain pointers and sizes only.
 *
 * Reference interfaces replaced (KAPSEL file:line):
 *   kh_step            Time_evolution_hydro (sp_3d_ns.cxx:41-194), the
 *                      function-pointer seam `Time_evolution` (sp_3d_ns.cxx:10);
 *                      after kh_set_shear, Time_evolution_hydro_OBL
 *                      (sp_3d_ns.cxx:328-433), bound at sp_3d_ns.cxx:704-707
 *   kh_init_fluid      Init_zeta_k + initial constraint block
 *                      (init_fluid.cxx:10-25, sp_3d_ns.cxx:809-826)
 *   kh_A2a_k/kh_A_k2a  A2a_k / A_k2a (fft_wrapper.h:641-681)
 *   kh_apply           the individual functions below Time_evolution_hydro,
 *                      listed with enum kh_fn
 *   kh_run_create      Init_fft (fft_wrapper.cxx:153-186), Mem_alloc_var
 *                      (sp_3d_ns.cxx:588-639), Particle_domain
 *                      (profile.cxx:22-65), the derived constants of
 *                      Set_global_parameters (input.cxx:424-600)
 */
#ifndef KAPSEL_HIP_H
#define KAPSEL_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "kapsel_particle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum kh_status {
    KH_OK              = 0,
    KH_ERR_ARG         = -1, /* invalid argument or parameter */
    KH_ERR_HIP         = -2, /* HIP runtime error */
    KH_ERR_NOMEM       = -3, /* device allocation failed */
    KH_ERR_UNSUPPORTED = -4, /* size or option this library does not provide */
    KH_ERR_STATE       = -5, /* call made in the wrong state */
    KH_ERR_NONFINITE   = -6  /* non-finite particle state detected */
} kh_status;

/* enum JP (input.h:49) */
enum { KH_JP_MOTOR = 0, KH_JP_SLIP = 1, KH_JP_OBSTACLE = 2, KH_JP_NONE = 3 };
/* enum JAX (input.h:48) */
enum { KH_JAX_X = 0, KH_JAX_Y = 1, KH_JAX_Z = 2, KH_JAX_NONE = 3 };
/* precision of the field path */
enum { KH_FP64 = 0, KH_MIXED = 1 };

/* One particle species: object_type.spherical_particle.Particle_spec[i]
 * (input.cxx:1400-1470). */
typedef struct kh_species {
    double mass_ratio;      /* MASS_RATIOS[i]: rho_particle / rho            */
    int    propulsion;      /* janus_propulsion[i] (KH_JP_*)                   */
    int    janus_axis;      /* janus_axis[i] (KH_JAX_*)                        */
    double janus_force[3];  /* body-frame motor force                          */
    double janus_torque[3]; /* body-frame motor torque                         */
    double slip_vel;        /* janus_slip_vel[i] = B1                          */
    double slip_mode;       /* janus_slip_mode[i] = B2/B1 as given in the input;
                               halved internally (input.cxx:581-585)           */
    double rotlet_c1;       /* janus_rotlet_C1[i]                              */
    double rotlet_c2;       /* janus_rotlet_dipole_C2[i]                       */
} kh_species;

/* Physical and numerical parameters, in the reference's input units
 * (input.cxx:685-2557); the derived constants are recomputed as
 * Set_global_parameters does (input.cxx:424-600). */
typedef struct kh_params {
    int    nx, ny, nz;  /* grid points per axis (mesh.NPX.. as 1<<NP); sizes 2^a 3^b, 8..512 */
    double dx;          /* DX */
    double rho, eta;    /* RHO, ETA */
    double kbt;         /* kBT; thermal noise on when > 0 */
    double alpha_v, alpha_o;
    double dt;          /* time_increment.manual.delta_t */
    double a, a_xi;     /* A, A_XI (grid units): RADIUS = A*DX, XI = A_XI*DX */
    double g;           /* gravity.G */
    int    g_dir;       /* gravity.G_direction: 0,1,2 for -X,-Y,-Z */
    int    rotation;    /* switch.ROTATION */
    int    fix_cell[3]; /* switch.FIX_CELL.{x,y,z} */
    int    lj_truncate; /* switch.LJ_truncate: 1 ON, 0 OFF, -1 NONE */
    int    lj_powers;   /* LJ_powers: 0 "12:6", 1 "24:12", 2 "36:18" */
    double epsilon;     /* EPSILON */
    int    lj_cell_compat; /* 1: count pairs with the multiplicity of the
                              reference's 16-cell link list (md_force.cxx:14,
                              56-101), which double-counts when N/16 == 2;
                              0: count every pair once */
    int               nspecies;
    const kh_species *species; /* nspecies entries, copied at kh_run_create */
    int    slip_iter_max;      /* switch.SLIP_iter (MAX_SLIP_ITER) */
    double slip_tol;           /* switch.SLIP_tol (MAX_SLIP_TOL) */
    int    precision;          /* KH_FP64 or KH_MIXED */
    uint64_t seed;             /* thermal-noise stream key */
    uint32_t stream_id;        /* second key word: independent runs, one seed */
} kh_params;

typedef struct kh_run kh_run;

/* Fields that can be read or written in the reference's real-space layout
 * (NX*NY*NZ_ doubles, NZ_ = 2*(NZ/2+1), padding columns ignored). */
enum kh_field {
    KH_FIELD_U0 = 0, KH_FIELD_U1, KH_FIELD_U2,    /* u   (fft_wrapper.cxx:20) */
    KH_FIELD_UP0, KH_FIELD_UP1, KH_FIELD_UP2,     /* up  (fft_wrapper.cxx:20) */
    KH_FIELD_PHI, KH_FIELD_PHI_SUM,               /* phi, phi_sum             */
    KH_FIELD_F0, KH_FIELD_F1, KH_FIELD_F2,        /* f_particle (f_particle.cxx:10) */
    KH_FIELD_COUNT
};

/* Single functions of the step, applied to the run's device state, for
 * per-function parity with the reference. */
enum kh_fn {
    KH_FN_NS_SOLVER = 0,      /* NS_solver_slavedEuler        fluid_solver.cxx:348-374 */
    KH_FN_ZETA_K2U,           /* Zeta_k2u (zeta,uk_dc -> u)   operate_omega.h:134-137 */
    KH_FN_U2ZETA_K,           /* U2zeta_k (u -> zeta,uk_dc)   operate_omega.h:208-211 */
    KH_FN_POSITION_EULER,     /* MD_solver_position_Euler     particle_solver.cxx:39-61 */
    KH_FN_POSITION_AB2,       /* MD_solver_position_AB2       particle_solver.cxx:63-85 */
    KH_FN_MAKE_PHI,           /* Reset_phi x2 + Make_phi_particle_sum  make_phi.cxx:214-260 */
    KH_FN_F_HYDRO,            /* Calc_f_hydro_correct_precision        md_force.cxx:383-506 */
    KH_FN_VELOCITY_EULER,     /* MD_solver_velocity_Euler     particle_solver.cxx:103-142 */
    KH_FN_VELOCITY_AB2,       /* MD_solver_velocity_AB2_hydro particle_solver.cxx:239-281 */
    KH_FN_THERMOSTAT,         /* Add_random_force_thermostat  fluct.cxx:10-100 */
    KH_FN_MAKE_UP,            /* Reset_u + Make_u_particle_sum         make_phi.cxx:313-354 */
    KH_FN_F_PARTICLE_SOLE,    /* Make_f_particle_dt_sole      f_particle.h:60-67 */
    KH_FN_ADD_F_PARTICLE,     /* Add_f_particle (u += f)      f_particle.h:68-84 */
    KH_FN_SLIP_MOMENTUM,      /* Make_particle_momentum_factor (u)     operate_surface.cxx:12-153 */
    KH_FN_SLIP_SCATTER,       /* Reset_u + Make_force_u_slip_particle  operate_surface.cxx:200-363 */
    KH_FN_SLIP_ITERATION,     /* the slip loop + Swap_mem(u, up)       sp_3d_ns.cxx:121-161 */
    /* sheared runs (kh_set_shear), at the run's strain and rate; on such a
       run the particle functions above take the Lees-Edwards geometry */
    KH_FN_OBL_NS_SOLVER,      /* NS_solver_slavedEuler_Shear_OBL fluid_solver.cxx:411-437 */
    KH_FN_OBL_ZETA_K2U,       /* Zeta_k2u_k_OBL + U_k2u        fft_wrapper.cxx:690-768 */
    KH_FN_OBL_U2ZETA_K,       /* U2u_k + Solenoidal_uk_OBL + U_k2zeta_k_OBL (sp_3d_ns.cxx:428-431) */
    KH_FN_OBL_OBLIQUE2U,      /* Transform_obl_u(u, oblique2cartesian) fft_wrapper.h:480-505 */
    KH_FN_COUNT
};

/* -------- lifecycle -------- */
int kh_run_create(const kh_params *params, int nparticles, kh_run **out);
int kh_run_destroy(kh_run *run);

/* -------- state transfer (host <-> HBM) -------- */
/* Particles in KAPSEL's own array-of-structs (variable.h:27-77). */
int kh_set_particles(kh_run *run, const Particle *p, int n);
int kh_get_particles(kh_run *run, Particle *p, int n);
/* zeta in the reference spectral layout (NX*NY*NZ_ doubles each,
 * (re,im) of mode (kx,ky,kz) at [i][j][2kz]); uk_dc[3]. */
int kh_set_fluid(kh_run *run, const double *zeta0, const double *zeta1, const double uk_dc[3]);
int kh_get_fluid(kh_run *run, double *zeta0, double *zeta1, double uk_dc[3]);
int kh_set_time(kh_run *run, const CTime *jikan);
int kh_get_time(kh_run *run, CTime *jikan);
int kh_get_field(kh_run *run, int field, double *out);
int kh_set_field(kh_run *run, int field, const double *in);

/* -------- the step -------- */
/* Fluid at rest plus the initial rigidity constraint (sp_3d_ns.cxx:809-826). */
int kh_init_fluid(kh_run *run);
/* nsteps calls of Time_evolution_hydro, each followed by ts += 1 and
 * time += dt (sp_3d_ns.cxx:857, 954-955). Enqueued on the run's stream;
 * returns once enqueued unless an error is detected. */
int kh_step(kh_run *run, int nsteps);
/* Kernel sequence kh_step uses. KH_PATH_FUSED (default): the dealiasing-
 * pruned, fused pipeline. KH_PATH_REFERENCE_ORDER: one kernel sequence per
 * reference function, in the reference's order (20 full 3D FFTs per step).
 * Both give the reference's step to round-off. */
enum { KH_PATH_FUSED = 0, KH_PATH_REFERENCE_ORDER = 1 };
int kh_set_step_path(kh_run *run, int path);

/* Squirmer runs: the slip iterations (Time_evolution_hydro's loop,
 * sp_3d_ns.cxx:121-161) run so far in total and the number of slip loops
 * (steps) they were spread over, since the run was created. A diagnostic
 * with no counterpart in the reference. */
int kh_get_slip_iterations(kh_run *run, long long *iterations, long long *loops);
/* Lees-Edwards shear (SW_EQ == Shear_Navier_Stokes_Lees_Edwards): from the
 * next step on kh_step runs Time_evolution_hydro_OBL (sp_3d_ns.cxx:328-433),
 * the sheared step on an oblique grid that follows the mean flow
 * rate*(y - LY/2): DC at shear_rate (ac = 0) or AC shear_rate*cos(2 pi
 * frequency t) (input.cxx:770-792). Particle positions and velocities stay
 * Cartesian (lab frame). LJ pairs are counted once and the LJ diameter is
 * min((SIGMA + XI)/2^(1/2n), SIGMA) as the reference's sheared runs take it
 * (input.cxx:549-566). Not for squirmer (slip) runs. The strain
 * degree_oblique (in (-NX/NY, NX/NY)) and Shear_rate_eff are run state,
 * saved and restored like zeta (resume.cxx:72, 566-580). */
int kh_set_shear(kh_run *run, double shear_rate, int ac, double frequency);
int kh_get_shear_state(kh_run *run, double *degree_oblique, double *shear_rate_eff);
int kh_set_shear_state(kh_run *run, double degree_oblique, double shear_rate_eff);
/* The record Mean_shear_stress prints after every sheared step
 * (sp_3d_ns.cxx:962-964, sp_3d_ns.h:247-330), kept on the device:
 * time, shear_rate, degree_oblique, strain (Shear_strain_realized),
 * lj_dev_stress, shear_stress_old, shear_stress_new (the Hydro_force sums of
 * md_force.cxx:508-745), reynolds_stress (Calc_Reynolds_shear_stress,
 * fluid_solver.h:165-183), fluid_stress (ETA * rate), interfacial_stress
 * (0), apparent_stress, viscosity. kh_get_shear_record copies the last
 * step's; kh_set_shear_history keeps the records of the next nsteps steps
 * (0 stops), kh_get_shear_history copies them (nsteps x KH_SHEAR_RECORD). */
#define KH_SHEAR_RECORD 12
int kh_get_shear_record(kh_run *run, double out[KH_SHEAR_RECORD]);
int kh_set_shear_history(kh_run *run, int nsteps);
int kh_get_shear_history(kh_run *run, double *out, int nsteps);
/* Rigid composite bodies (SW_PT == rigid; rigid.h, make_phi.cxx:589-801).
 * Body k is made of the particles (beads) cum[k] .. cum[k+1]-1
 * (Rigid_Particle_Cumul), its species that of its first bead; the beads'
 * x_nopbc must hold coordinates that place each body in one piece (the
 * reference reads them unwrapped, init_particle.cxx:455-487), x the same
 * wrapped into the box. free_flags[k]: 1 "free", 0 "fix" (Rigid_motion;
 * NULL: all free). v0, w0: initial body velocities, nb x 3, space frame
 * (NULL: at rest, as the reference starts them). Call after
 * kh_set_particles and before kh_init_fluid: it runs the reference's
 * initial block -- centres and masses from the phi-weighted bead volume,
 * inertia tensors, principal frames (the bead orientations q), bead
 * velocities -- and from then on kh_step moves bodies: per-body sums of the
 * beads' hydrodynamic force and torque about the centre, LJ between beads of
 * different bodies only (with lj_cell_compat the reference's pair
 * bookkeeping, md_force.cxx:134-153), gravity on the body volume, Euler's
 * equations in the body frame (calc_Rigid_VOGs, rigid.h:467-559),
 * positions and orientations (solver_Rigid_Position, rigid.h:333-392).
 * Not for kT > 0, slip species, sheared runs or ROTATION off
 * (KH_ERR_UNSUPPORTED). */
#define KH_RIGID_WORDS 32
int kh_set_rigid(kh_run *run, int nbodies, const int *cum, const int *free_flags, const double *v0,
                 const double *w0);
/* Per body, KH_RIGID_WORDS doubles: xG[3], xG_nopbc[3], vG[3], omegaG[3],
 * the previous step's hydrodynamic force[3] and torque[3] and LJ + gravity
 * force[3] and torque[3] (the RigidParticles members of Output_udf,
 * output.cxx:324-373), mass, principal moments[3], orientation q[4] (q0
 * scalar). */
int kh_get_rigid(kh_run *run, double *out, int nbodies);

/* Per-step particle record kept on the device: from the next step on, the
 * (x[3], v[3]) of every particle after each of the next nsteps steps
 * (nsteps x np x 6 doubles; 0 stops recording). No counterpart in the
 * reference step; it feeds trajectory output (Output_udf's per-frame
 * particle records, output.cxx:268-375) without a host copy per step. */
int kh_set_trajectory(kh_run *run, int nsteps);
int kh_get_trajectory(kh_run *run, double *out, int nsteps);
/* Output_udf frames (output.cxx:268-375, written at ts % GTS == 0 before the
 * step of that ts, sp_3d_ns.cxx:863-950) kept on the device: frames at the
 * next nframes step counts ts >= now with ts % every == 0 (the current state
 * when ts % every == 0 already). Frame k is 1 + KH_FRAME_WORDS * np doubles:
 * ts, then per particle the record's members in its order: R[3], R_raw[3],
 * v[3], q[4] (q0 scalar), omega[3], f_hydro[3], torque_hydro[3], f_r[3],
 * torque_r[3] (0), f_slip[3], torque_slip[3]; the forces and torques are the
 * previous step's (the *_previous members the reference writes). The record's
 * constant E = 1.0 is not stored. nframes = 0 stops recording. */
#define KH_FRAME_WORDS 34
int kh_set_output(kh_run *run, int nframes, int every);
int kh_get_output(kh_run *run, double *out, int nframes);
/* 1 when all work enqueued on the run (kh_step is asynchronous) has
 * completed, 0 while some is pending, a negative status on a device error. */
int kh_query(kh_run *run);
/* Device bytes kh_run_create would allocate for these parameters. */
long long kh_run_bytes(const kh_params *params, int nparticles);

/* Steady-state steps (ts >= 1) of the fused path run as one captured HIP
 * graph by default (captured on first use); 0 turns that off. Kernel
 * timing (kh_timing_enable) also runs steps uncaptured. */
int kh_set_graph(kh_run *run, int on);
/* One function of the step on the device state (enum kh_fn). */
int kh_apply(kh_run *run, int fn);
/* Wait for the run's stream; reports errors raised by enqueued work. After
   sheared steps it first forms the vorticity modes beyond the 2/3 rule that
   they leave to be formed, so the device state is complete when it returns. */
int kh_sync(kh_run *run);

/* -------- standalone transforms (A2a_k / A_k2a semantics) --------
 * In place on a host array in the reference layout (NX*NY*NZ_ doubles).
 * Forward: unnormalised sum a(x) exp(-i k.x). Inverse: normalised by
 * 1/(NX*NY*NZ); the kz=0 and kz=NZ/2 planes are taken as Hermitian. */
int kh_A2a_k(kh_run *run, double *a);
int kh_A_k2a(kh_run *run, double *a);

/* kh_A2a_k without a run and without any device allocation. No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). FP64 forward transform, in place, of a host array in the
 * reference layout (NX*NY*NZ_ doubles, NZ_ = NZ + 2), the result as
 * kh_A2a_k gives it. Its staging (the real field and its spectrum, FP64,
 * spectral rows padded to 128 bytes) lives in a fixed device array of
 * KH_NOALLOC_MAX_BYTES that each GPU holds until the process ends; larger
 * grids (more than 128^3 or so) return KH_ERR_UNSUPPORTED. Runs on the
 * current device's null stream and returns when the result is in `a`.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race. */
#define KH_NOALLOC_MAX_BYTES (64ll << 20)
int kh_A2a_k_noalloc(double *a, int nx, int ny, int nz);


/* -------- measurement -------- */
/* When enabled, every kernel launch of the run is bracketed by HIP events
 * on the run's stream; kh_timing_get returns the launch count and the summed
 * event time of kernel `name` since the last reset. */
int kh_timing_enable(kh_run *run, int on);
int kh_timing_reset(kh_run *run);
int kh_timing_get(kh_run *run, const char *name, int *launches, double *total_ms);
/* Number of distinct kernels timed and the name of the i-th. */
int kh_timing_count(kh_run *run);
const char *kh_timing_name(kh_run *run, int i);

const char *kh_strerror(int status);
const char *kh_last_error(void);

/* ---------------------------------------------------------------------
 * Serving: continuous batching of independent runs on one GPU. No
 * counterpart in the reference (one simulation per process there); it
 * runs many SPM simulations (Time_evolution_hydro steps, as kh_step) side by
 * side. Runs are admitted in arrival order while fewer than max_active are
 * resident and their kh_run_bytes() fit the budget (hbm_budget bytes, or
 * hbm_fraction of the device memory free at creation); each stays on the
 * device until its nsteps are done, then its trajectory (as
 * kh_get_trajectory) is copied to traj_out (nsteps x np x 6 doubles, may be
 * NULL) and its memory is released. Latency = submission to that copy.
 * Runs of plain spheres on the same grid (size class: grid, precision, dx,
 * rho, eta, dt, FIX_CELL, a, a_xi) with at most KH_BATCH_NP_CAP particles
 * share a batch of batch_slots device arenas: after a run's first step, one
 * launch per fluid pass and one captured graph step every run of the batch
 * together (its particle launches stay its own); the batch's arenas count
 * against the budget while any of them is in use.
 * Submit/poll/wait/stats may be called from any thread. */
typedef struct kh_server kh_server;
typedef struct {
    int       max_active;    /* runs resident at once (>= 1)                       */
    int       chunk_steps;   /* steps enqueued per run per scheduler visit (>= 1)  */
    double    hbm_fraction;  /* share of free device memory for runs, (0, 1]      */
    long long hbm_budget;    /* bytes; > 0 overrides hbm_fraction                  */
    int       batch_slots;   /* runs per batch of one size class; < 2: every run steps alone */
} kh_server_config;
#define KH_BATCH_NP_CAP 64
typedef struct {
    long long submitted, completed, failed, queued, active;
    int       peak_active;
    long long peak_bytes, budget;
    double    p50_ms, p99_ms, mean_ms, max_ms;  /* submission -> trajectory copied  */
    double    wall_s;                          /* first submission -> last completion */
    double    runs_per_s, steps_per_s, grid_point_steps_per_s;
} kh_server_stats;
int kh_server_create(const kh_server_config *cfg, kh_server **out);
int kh_server_submit(kh_server *srv, const kh_params *params, const Particle *particles, int nparticles,
                     int nsteps, double *traj_out, long long *id);
/* As kh_server_submit, and the run's Output_udf frames (output.cxx:268-375,
 * as kh_set_output records them) at ts = 0, frame_every, 2 frame_every, ...
 * up to nsteps are copied to frames_out at completion:
 * (nsteps / frame_every + 1) frames of 1 + KH_FRAME_WORDS * np doubles.
 * The reference writes a frame every GTS steps of its loop
 * (sp_3d_ns.cxx:863-950); frame_every is that GTS. */
int kh_server_submit_output(kh_server *srv, const kh_params *params, const Particle *particles, int nparticles,
                            int nsteps, double *traj_out, int frame_every, double *frames_out, long long *id);
/* 1 done, 0 pending, < 0 the run's failure status (kh_last_error has it) */
int kh_server_poll(kh_server *srv, long long id);
/* as poll, blocking up to timeout_s seconds (0 on timeout) */
int kh_server_wait(kh_server *srv, long long id, double timeout_s);
/* seconds queued and seconds resident of a completed run */
int kh_server_run_times(kh_server *srv, long long id, double *queued_s, double *run_s);
int kh_server_get_stats(kh_server *srv, kh_server_stats *out);
/* finishes every submitted run, then frees the server */
int kh_server_destroy(kh_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* KAPSEL_HIP_H */
