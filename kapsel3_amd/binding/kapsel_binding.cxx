// kapsel_binding.cxx -- see kapsel_binding.h. Compiled into KAPSEL (C++11,
// with its variable.h on the include path); the device side is
// libkapsel_hip. Errors follow the reference's convention: a message on
// stderr and exit_job(EXIT_FAILURE) (macro.h:178).
#include <stdio.h>
#include <stdlib.h>

#include "variable.h"
#include "kapsel_binding.h"

namespace {

kh_run *g_run = nullptr;
int     g_np  = 0;

void check(int s, const char *what) {
    if (s != KH_OK) {
        fprintf(stderr, "kapsel_hip: %s: %s: %s\n", what, kh_strerror(s), kh_last_error());
        exit_job(EXIT_FAILURE);
    }
}
kh_run *run() {
    if (!g_run) {
        fprintf(stderr, "kapsel_hip: kh_binding_init was not called\n");
        exit_job(EXIT_FAILURE);
    }
    return g_run;
}

}  // namespace

void kh_binding_init(const kh_params *params, int nparticles, Particle *p, double **zeta, double uk_dc[DIM]) {
    kh_binding_finish();
    check(kh_run_create(params, nparticles, &g_run), "kh_run_create");
    g_np = nparticles;
    check(kh_set_particles(g_run, p, nparticles), "kh_set_particles");
    check(kh_set_fluid(g_run, zeta[0], zeta[1], uk_dc), "kh_set_fluid");
}

void kh_binding_finish() {
    if (g_run) check(kh_run_destroy(g_run), "kh_run_destroy");
    g_run = nullptr;
}

void Time_evolution_hydro_hip(double **zeta, double uk_dc[DIM], double **f, Particle *p, CTime &jikan) {
    (void)zeta;
    (void)uk_dc;
    (void)f;
    (void)p;
    check(kh_set_time(run(), &jikan), "kh_set_time");
    check(kh_step(run(), 1), "kh_step");
}

void kh_binding_download(double **zeta, double uk_dc[DIM], Particle *p) {
    check(kh_get_fluid(run(), zeta[0], zeta[1], uk_dc), "kh_get_fluid");
    check(kh_get_particles(run(), p, g_np), "kh_get_particles");
}

void A2a_k_hip(double *a) { check(kh_A2a_k(run(), a), "kh_A2a_k"); }
void A_k2a_hip(double *a) { check(kh_A_k2a(run(), a), "kh_A_k2a"); }

void NS_solver_slavedEuler_hip(double **zeta, const CTime &jikan, double uk_dc[DIM]) {
    check(kh_set_fluid(run(), zeta[0], zeta[1], uk_dc), "kh_set_fluid");
    check(kh_set_time(run(), &jikan), "kh_set_time");
    check(kh_apply(run(), KH_FN_NS_SOLVER), "NS_solver_slavedEuler");
    check(kh_get_fluid(run(), zeta[0], zeta[1], uk_dc), "kh_get_fluid");
}

void Make_phi_particle_sum_hip(double *phi, double *phi_sum, Particle *p) {
    check(kh_set_particles(run(), p, g_np), "kh_set_particles");
    check(kh_apply(run(), KH_FN_MAKE_PHI), "Make_phi_particle_sum");
    check(kh_get_field(run(), KH_FIELD_PHI, phi), "phi");
    check(kh_get_field(run(), KH_FIELD_PHI_SUM, phi_sum), "phi_sum");
}

void Make_u_particle_sum_hip(double **up, double const *phi_sum, Particle *p) {
    check(kh_set_particles(run(), p, g_np), "kh_set_particles");
    check(kh_set_field(run(), KH_FIELD_PHI_SUM, phi_sum), "phi_sum");
    check(kh_apply(run(), KH_FN_MAKE_UP), "Make_u_particle_sum");
    for (int d = 0; d < DIM; d++) check(kh_get_field(run(), KH_FIELD_UP0 + d, up[d]), "up");
}

void Calc_f_hydro_correct_precision_hip(Particle *p, double const *phi_sum, double const *const *u,
                                        const CTime &jikan) {
    check(kh_set_time(run(), &jikan), "kh_set_time");
    check(kh_set_particles(run(), p, g_np), "kh_set_particles");
    check(kh_set_field(run(), KH_FIELD_PHI_SUM, phi_sum), "phi_sum");
    for (int d = 0; d < DIM; d++) check(kh_set_field(run(), KH_FIELD_U0 + d, u[d]), "u");
    check(kh_apply(run(), KH_FN_F_HYDRO), "Calc_f_hydro_correct_precision");
    check(kh_get_particles(run(), p, g_np), "kh_get_particles");
}
