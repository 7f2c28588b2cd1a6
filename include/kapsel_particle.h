/*
 * kapsel_particle.h -- host-side data contract shared with KAPSEL's step.
 *
 * Mirrors, field for field, the structs the step is called with, so that an
 * array of them can be handed across the C-ABI unchanged:
 *   - CTime      : variable.h:18-25   (48 bytes)
 *   - quaternion : quaternion.h       (32 bytes: scalar part, vector part)
 *   - Particle   : variable.h:27-77   (1048 bytes, array-of-structs)
 * The offsets asserted below are the ones KAPSEL's own compiler lays out
 * (x@8, v@80, f_hydro@152, omega@368, q@728); the library reads and writes
 * only the members the SPM step touches.
 *
 * Valid C99 and C++11. When KAPSEL's own variable.h / quaternion.h have
 * been included first (include guards VARIABLE_H / QUATERNION_H), their
 * definitions are used and only the layout checks below apply.
 */
#ifndef KAPSEL_PARTICLE_H
#define KAPSEL_PARTICLE_H

#include <stddef.h>

#ifndef DIM
#define DIM 3
#endif

#ifndef VARIABLE_H
typedef struct CTime {
    int    ts;        /* time step */
    double time;      /* time */
    double dt_fluid;  /* time increment */
    double hdt_fluid; /* dt/2 */
    double dt_md;     /* time increment */
    double hdt_md;    /* dt/2 */
} CTime;

#endif /* VARIABLE_H */

#ifndef QUATERNION_H
typedef struct quaternion {
    double s;      /* scalar part */
    double v[DIM]; /* vector part */
} quaternion;
#endif /* QUATERNION_H */

#ifndef VARIABLE_H
typedef struct Particle {
    int    spec;
    double x[DIM];
    double x_previous[DIM];
    double x_nopbc[DIM];

    double v[DIM];
    double v_old[DIM];
    double v_slip[DIM];

    double f_hydro[DIM];
    double f_hydro_previous[DIM];
    double f_hydro1[DIM];
    double f_slip[DIM];
    double f_slip_previous[DIM];

    double fr[DIM];
    double fr_previous[DIM];
    double torque_r[DIM];
    double torque_r_previous[DIM];

    double omega[DIM];
    double omega_old[DIM];
    double omega_slip[DIM];

    double torque_hydro[DIM];
    double torque_hydro_previous[DIM];
    double torque_hydro1[DIM];
    double torque_slip[DIM];
    double torque_slip_previous[DIM];

    double momentum_depend_fr[DIM];

    double QR[DIM][DIM];
    double QR_old[DIM][DIM];

    quaternion q;
    quaternion q_old;

    double mass;
    double mass_center[DIM];
    double inertia[DIM][DIM];

    double surface_mass;
    double surface_mass_center[DIM];
    double surface_inertia[DIM][DIM];

    double surface_dv[DIM];
    double surface_dw[DIM];
} Particle;
#endif /* VARIABLE_H */

#ifdef __cplusplus
#define KH_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define KH_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

KH_STATIC_ASSERT(sizeof(CTime) == 48, "CTime layout");
KH_STATIC_ASSERT(sizeof(quaternion) == 32, "quaternion layout");
KH_STATIC_ASSERT(sizeof(Particle) == 1048, "Particle layout");
KH_STATIC_ASSERT(offsetof(Particle, x) == 8, "Particle.x");
KH_STATIC_ASSERT(offsetof(Particle, v) == 80, "Particle.v");
KH_STATIC_ASSERT(offsetof(Particle, f_hydro) == 152, "Particle.f_hydro");
KH_STATIC_ASSERT(offsetof(Particle, omega) == 368, "Particle.omega");
KH_STATIC_ASSERT(offsetof(Particle, q) == 728, "Particle.q");

#endif /* KAPSEL_PARTICLE_H */
