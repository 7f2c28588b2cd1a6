// kapsel_binding.h -- the KAPSEL side of the MI355X SPM step: functions with
// the reference's own signatures that run on the device through the C ABI
// of kapsel_hip.h. Include after KAPSEL's variable.h (Particle, CTime) and
// link with libkapsel_hip.
#ifndef KAPSEL_BINDING_H
#define KAPSEL_BINDING_H

#include <vector>

#include "kapsel_hip.h"

// kh_params from KAPSEL's globals (kapsel_binding_globals.cxx, needs input.h);
// `spec` holds the species array the parameters point to.
kh_params kh_params_from_globals(std::vector<kh_species> &spec, int precision);

// Creates the device run from parameters (see kh_params_from_globals in
// kapsel_binding_globals.cxx) and uploads the particles and the fluid.
void kh_binding_init(const kh_params *params, int nparticles, Particle *p, double **zeta, double uk_dc[DIM]);
void kh_binding_finish();

// The seam: same signature as `Time_evolution` (sp_3d_ns.cxx:10) and
// Time_evolution_hydro (sp_3d_ns.cxx:41-194); after kh_set_shear on the run
// it is Time_evolution_hydro_OBL (sp_3d_ns.cxx:328-433). `f` is unused: the
// device holds its own work fields.
void Time_evolution_hydro_hip(double **zeta, double uk_dc[DIM], double **f, Particle *p, CTime &jikan);
// Host copies of the state for Output_* and Save_Restart_udf
// (sp_3d_ns.cxx:863-975).
void kh_binding_download(double **zeta, double uk_dc[DIM], Particle *p);

// Single functions of the step, each on host arrays in KAPSEL's layout:
// upload, one kh_apply, download.
void A2a_k_hip(double *a);                                       // A2a_k  (fft_wrapper.h:641-662)
void A_k2a_hip(double *a);                                       // A_k2a  (fft_wrapper.h:663-681)
void NS_solver_slavedEuler_hip(double **zeta, const CTime &jikan, double uk_dc[DIM]);  // fluid_solver.h:77-82
void Make_phi_particle_sum_hip(double *phi, double *phi_sum, Particle *p);              // make_phi.h:50-52
void Make_u_particle_sum_hip(double **up, double const *phi_sum, Particle *p);          // make_phi.h:53-55
void Calc_f_hydro_correct_precision_hip(Particle *p, double const *phi_sum, double const *const *u,
                                        const CTime &jikan);                           // md_force.h:53

#endif
