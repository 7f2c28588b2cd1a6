// kh_launch.h -- host launchers of the SPM kernels (one stream argument each).
#pragma once

#include "kh_common.h"
#include "kh_fft.h"
#include "kh_particle.h"

namespace kh {

// kh_spectral.hip
template <typename T>
hipError_t zeta_to_u(const Grid &g, const cpx<T> *z0, const cpx<T> *z1, cpx<T> *u0, cpx<T> *u1, cpx<T> *u2,
                     double *ukdc, int fixmask, int truncate, hipStream_t s);
template <typename T>
hipError_t products(long long n, const T *u0, const T *u1, const T *u2, T *const p[5], hipStream_t s);
template <typename T>
hipError_t adv_ns(const Grid &g, cpx<T> *const P[5], cpx<T> *z0, cpx<T> *z1, double dmy0, double dmy1,
                  hipStream_t s);
template <typename T>
hipError_t solenoidal(const Grid &g, cpx<T> *u0, cpx<T> *u1, cpx<T> *u2, hipStream_t s);
template <typename T>
hipError_t u_to_zeta(const Grid &g, const cpx<T> *u0, const cpx<T> *u1, const cpx<T> *u2, cpx<T> *z0, cpx<T> *z1,
                     double *ukdc, hipStream_t s);
template <typename T>
hipError_t f_nonsole(long long n, T *const up[3], const T *phisum, T *const u[3], T *const f[3], hipStream_t s);
template <typename T>
hipError_t add3(long long n, T *const u[3], T *const f[3], hipStream_t s);
template <typename T>
hipError_t fill(long long n, T *p, T v, hipStream_t s, const Grid *batch = nullptr);

// kh_shear.hip (the Lees-Edwards step; gd = strain of the oblique grid)
template <typename T>
hipError_t obl_zeta_to_u(const Grid &g, cpx<T> *z0, cpx<T> *z1, cpx<T> *u0, cpx<T> *u1, cpx<T> *u2, cpx<T> *w1,
                         double *ukdc, int fixmask, int truncate, double gd, hipStream_t s);
template <typename T>
hipError_t obl_zeta_to_omega(const Grid &g, const cpx<T> *z0, const cpx<T> *z1, cpx<T> *const o[3], hipStream_t s);
template <typename T>
hipError_t obl_cross(long long n, T *const u[3], T *const w[3], hipStream_t s);
template <typename T>
hipError_t obl_adv_ns(const Grid &g, cpx<T> *const F[3], const cpx<T> *w1, cpx<T> *z0, cpx<T> *z1, double rate,
                      double gd, double dmy0, double dmy1, hipStream_t s);
template <typename T>
hipError_t obl_remap(const Grid &g, T *const u[3], T *const o[3], double delta, hipStream_t s);
template <typename T>
hipError_t obl_to_cart(const Grid &g, T *const u[3], T *const c[3], double gd, double rate, const T *mask,
                       double *acc, double dx, hipStream_t s);
long long obl_to_cart_blocks(const Grid &g);
template <typename T>
hipError_t obl_add_f(const Grid &g, T *const up[3], const T *phisum, T *const c[3], T *const u[3], double gd,
                     hipStream_t s);
template <typename T>
hipError_t obl_u_to_zeta(const Grid &g, const cpx<T> *u0, const cpx<T> *u1, const cpx<T> *u2, cpx<T> *z0, cpx<T> *z1,
                         double *ukdc, double gd, hipStream_t s, int kept = 0);

// kh_particle.hip (PCache: per-particle cache of the Sekibun points with
// nonzero phi, filled by phi_scatter)
// move = 1: the particle is first advanced by the position integrator
template <typename T>
hipError_t phi_scatter(const PParams &pp, DevParticle *P, T *phisum, const PCache<T> &pc, int scatter, hipStream_t s,
                       int move = 0, int ab2 = 0, double dt = 0.0, double hdt = 0.0);
template <typename T>
hipError_t f_hydro(const PParams &pp, DevParticle *P, T *const u[3], const T *phisum, const PCache<T> &pc,
                   double coef, hipStream_t s);
template <typename T>
hipError_t up_scatter(const PParams &pp, const DevParticle *P, const T *phisum, const PCache<T> &pc,
                      T *const up[3], hipStream_t s);
hipError_t position(const PParams &pp, DevParticle *P, int ab2, double dt, double hdt, hipStream_t s);
hipError_t forces(const PParams &pp, DevParticle *P, hipStream_t s);
// rigid bodies: solver_Rigid_Position + update_Orientation +
// update_Particle_Configuration, and calc_Rigid_VOGs + set_Particle_Velocities
// + reset_Forces (one workgroup per body)
hipError_t rigid_position(const PParams &pp, DevParticle *P, const RigidArgs &ra, int ab2, double dt, double hdt,
                          hipStream_t s);
hipError_t rigid_velocity(const PParams &pp, DevParticle *P, const RigidArgs &ra, int ab2, double dt, double hdt,
                          hipStream_t s);
// noise = 1: the thermostat's kick follows the update in the same launch
hipError_t velocity(const PParams &pp, DevParticle *P, int ab2, double h, hipStream_t s, int noise = 0,
                    uint64_t seed = 0, uint32_t stream = 0, const int *d_ts = nullptr, double niv = 0.0,
                    double nio = 0.0);
hipError_t thermostat(const PParams &pp, DevParticle *P, uint64_t seed, uint32_t stream, const int *d_ts, double niv,
                      double nio, hipStream_t s);
hipError_t tick(int *d_ts, hipStream_t s, const Grid *batch = nullptr);
// the sheared step's stress record from the per-particle partials and the
// nacc Reynolds partials of k_obl_to_cart
hipError_t obl_record(const PParams &pp, const double *acc, long long nacc, const ShearRec &h, double *last,
                      double *hist, hipStream_t s);
hipError_t frame(const PParams &pp, const DevParticle *P, const int *d_ts, int ts0, int every, int nframes, double *out,
                 hipStream_t s);
hipError_t record(const PParams &pp, const DevParticle *P, const int *d_ts, int base, int nslots, double *out,
                  hipStream_t s);
// F_H gather, LJ + gravity, velocity (+ thermostat kick), u_p scatter in one
// launch (the fused step after the phi scatter)
template <typename T>
hipError_t particle_step(const PParams &pp, DevParticle *P, T *const u[3], const T *phisum, const PCache<T> &pc,
                         double coef, int ab2, double h, int noise, uint64_t seed, uint32_t stream, const int *d_ts,
                         double niv, double nio, T *const up[3], hipStream_t s);
// the phi scatter (with the position update), the fused particle step and
// the records of a batch of runs, grid z = slot (kh_batch; at most 256
// particles per run)
template <typename T>
hipError_t phi_scatter_batch(const SlotPart<T> *sp, const int *active, int nslots, int np_cap, double dt, double hdt,
                             hipStream_t s);
template <typename T>
hipError_t particle_step_batch(const SlotPart<T> *sp, const int *active, int nslots, int np_cap, double coef, double h,
                               hipStream_t s);
template <typename T>
hipError_t records_batch(const SlotPart<T> *sp, const int *active, int nslots, hipStream_t s);
// squirmer slip iteration (sp_3d_ns.cxx:121-161). A non-null `skip` makes
// a launch do nothing once *skip is nonzero: the loop decided on the device
// (slip_conv with a SlipLoop) launches every iteration up to the maximum,
// and those after convergence return at once.
template <typename T>
hipError_t slip_moments(const PParams &pp, DevParticle *P, T *const u[3], hipStream_t s);
template <typename T>
hipError_t slip_scatter(const PParams &pp, DevParticle *P, T *const u[3], T *const up[3], hipStream_t s,
                        const int *skip = nullptr);
template <typename T>
hipError_t f_slip(const PParams &pp, DevParticle *P, T *const sf[3], double coef, hipStream_t s,
                  const int *skip = nullptr);
hipError_t velocity_slip(const PParams &pp, DevParticle *P, int ab2, double h, hipStream_t s,
                         const int *skip = nullptr);
hipError_t reset_forces(const PParams &pp, DevParticle *P, hipStream_t s);
hipError_t slip_update(const PParams &pp, DevParticle *P, int iter, int restore, hipStream_t s,
                       const int *skip = nullptr);
// Device state of a slip loop decided on the device: flag[0] is set once the
// loop has converged or reached its maximum, flag[1] counts this step's
// iterations; stat[0] totals iterations and stat[1] completed loops.
struct SlipLoop {
    int                *flag;
    unsigned long long *stat;
    double              tol;
    int                 max_iter;
};
// eps = the convergence measure; with a non-null loop (its flag clear) the
// iteration is counted and the flag set when eps < tol or the count reaches
// the maximum
hipError_t slip_conv(const PParams &pp, const DevParticle *P, double *eps, hipStream_t s,
                     const SlipLoop *loop = nullptr);
// flag[0] = flag[1] = 0: a new loop
hipError_t slip_loop_begin(const SlipLoop &loop, hipStream_t s);

}  // namespace kh
