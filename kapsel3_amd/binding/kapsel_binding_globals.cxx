// kapsel_binding_globals.cxx -- kh_params from KAPSEL's globals (input.h,
// set by Gourmet_file_io / Set_global_parameters, input.cxx:424-2557), and
// the lines a maintainer changes in main (sp_3d_ns.cxx) to run the step on
// the GPU. Compiled into KAPSEL next to kapsel_binding.cxx.
#include <vector>

#include "input.h"
#include "kapsel_binding.h"

kh_params kh_params_from_globals(std::vector<kh_species> &spec, int precision) {
    spec.resize(Component_Number);
    for (int i = 0; i < Component_Number; i++) {
        kh_species &s = spec[i];
        s.mass_ratio  = MASS_RATIOS[i];
        s.propulsion  = janus_propulsion[i];  // enum JP has the values of KH_JP_*
        s.janus_axis  = janus_axis[i];        // enum JAX has the values of KH_JAX_*
        for (int d = 0; d < DIM; d++) {
            s.janus_force[d]  = janus_force[i][d];
            s.janus_torque[d] = janus_torque[i][d];
        }
        s.slip_vel  = janus_slip_vel[i];
        s.slip_mode = 2.0 * janus_slip_mode[i];  // kh_params takes B2/B1 as read; KAPSEL halved it
        s.rotlet_c1 = janus_rotlet_C1[i];
        s.rotlet_c2 = janus_rotlet_dipole_C2[i];
    }
    kh_params p = {};
    p.nx        = NX;
    p.ny        = NY;
    p.nz        = NZ;
    p.dx        = DX;
    p.rho       = RHO;
    p.eta       = ETA;
    p.kbt       = kBT;
    p.alpha_v   = alpha_v;
    p.alpha_o   = alpha_o;
    p.dt        = DT;
    p.a         = A;
    p.a_xi      = A_XI;
    p.g         = G;
    p.g_dir     = G_direction;
    p.rotation  = ROTATION;
    for (int d = 0; d < DIM; d++) p.fix_cell[d] = FIX_CELLxyz[d];
    p.lj_truncate    = LJ_truncate;
    p.lj_powers      = LJ_powers;
    p.epsilon        = EPSILON;
    p.lj_cell_compat = 1;  // the pair multiplicity of md_force.cxx:34-196
    p.nspecies       = Component_Number;
    p.species        = spec.data();
    p.slip_iter_max  = MAX_SLIP_ITER;
    p.slip_tol       = MAX_SLIP_TOL;
    p.precision      = precision;  // KH_FP64 or KH_MIXED
    p.seed           = 0;
    p.stream_id      = 0;
    return p;
}

/* In main (sp_3d_ns.cxx), after the initial block (809-826):

     std::vector<kh_species> spec;
     kh_params prm = kh_params_from_globals(spec, KH_FP64);
     kh_binding_init(&prm, Particle_Number, particles, zeta, uk_dc);
     Time_evolution = Time_evolution_hydro_hip;          // was Time_evolution_hydro (692-694)

   at the top of the `jikan.ts % GTS == 0` block (863) and before
   Save_Restart_udf (967-975):

     kh_binding_download(zeta, uk_dc, particles);

   and kh_binding_finish() before the end of main. */
