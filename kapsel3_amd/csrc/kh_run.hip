// kh_run.hip -- the C ABI of include/kapsel_hip.h: one simulation's state
// resident in HBM and the SPM step (Time_evolution_hydro,
// sp_3d_ns.cxx:41-194) enqueued on the run's own stream.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/kapsel_hip.h"
#include "kh_batch.h"
#include "kh_fused.h"
#include "kh_launch.h"
#include "kh_zwave.h"

using namespace kh;

namespace {

thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int status, const char *fmt, ...) {
    char    buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

#define KH_HIP(call)                                                                                          \
    do {                                                                                                      \
        hipError_t e_ = (call);                                                                               \
        if (e_ != hipSuccess) return fail(KH_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
    } while (0)

#define KH_TRY(call)                 \
    do {                             \
        int s_ = (call);             \
        if (s_ != KH_OK) return s_;  \
    } while (0)

constexpr double KH_PI     = 3.14159265358979323846;
constexpr double kT_snap_v = 1.439;  // input.cxx:309
constexpr double kT_snap_o = 1.22;   // input.cxx:311

struct KernelTimes {
    int    launches = 0;
    double ms       = 0.0;
};
struct PendingEvent {
    std::string name;
    hipEvent_t  a, b;
};

}  // namespace

// Sets the calling thread's kh_last_error() text (used by kh_server.hip).
int kh_set_error(int status, const char *msg) {
    g_last_error = msg;
    return status;
}

struct kh_run {
    kh_params               P;
    std::vector<kh_species> species;
    int                     device = 0;
    hipStream_t             stream = nullptr;
    hipStream_t             own_stream = nullptr;  // stream created with the run (stream differs while batched)
    // a run inside a batch arena (kh_batch): its device buffers are
    // sub-allocated from arena in a fixed order, slot * stride after the
    // batch's first arena
    char                   *arena      = nullptr;
    size_t                  arena_used = 0;
    int                     slot       = -1;
    bool                    batched    = false;
    Grid                    g;
    PParams                 pp;
    int                     mixed = 0;
    size_t                  esz   = 8;
    int                     np    = 0;
    // derived constants (Set_global_parameters, input.cxx:424-600)
    double radius, xi, hxi, rho, eta, nu, dx3, kbt, lj_cutoff, gravity_on_fluid;
    std::vector<double> mass, imass, moi, imoi;
    int    motor = 0, slip = 0, fixmask = 0;
    // device buffers
    void       *zeta[2]  = {nullptr, nullptr};
    void       *spec[5]  = {};
    void       *ureal[3] = {};
    void       *up[3]    = {};
    void       *work[5]  = {};
    void       *phisum   = nullptr;
    void       *occ      = nullptr;  // occupancy bits of the 4-point z vectors (Occ)
    int         occw     = 0;        // words per row
    double     *ukdc     = nullptr;
    DevParticle *dp      = nullptr;
    DevSpecies  *ds      = nullptr;
    int         *sek     = nullptr;
    void        *cphi    = nullptr;  // per-particle phi at each Sekibun point (T)
    int         *cidx    = nullptr;  // per-particle grid index of each Sekibun point
    int         *ccode   = nullptr;  // per-particle packed Sekibun offset of each cached point
    int         *ccnt    = nullptr;  // per-particle count of cached (nonzero-phi) points
    int          pcache_valid = 0;
    int         *d_ts         = nullptr;  // jikan.ts on the device (noise counter)
    double      *slipeps      = nullptr;  // slip convergence measure (1 double)
    double      *xs           = nullptr;  // compact particle positions (np x 4)
    bool         zclear       = false;    // the fused step's z_force leaves u_p and phi_sum zero
    bool         pclean       = false;    // u_p and phi_sum are zero now
    double      *slipdc       = nullptr;  // k = 0 mode of the projected slip field (3 doubles)
    int          slip_iters   = 0;        // iterations of the last slip loop decided on the host
    long long    slip_iters_host = 0, slip_loops_host = 0;  // totals of the loops decided on the host
    // the fused step's slip loop decided on the device (power-of-two NZ >= 32):
    // int flag[2] (converged, iterations this step), then unsigned long long
    // stat[2] (total iterations, loops) at byte 16
    bool         slip_dev     = false;
    int         *slipflag     = nullptr;
    double      *traj         = nullptr;  // per-step particle record (x, v) for traj_steps steps
    int          traj_steps   = 0;
    int          traj_base    = 0;        // ts when recording started
    double      *frames       = nullptr;  // Output_udf frames (kh_set_output)
    int          frame_n = 0, frame_every = 1, frame_ts0 = 0;
    // the steady-state step (ts >= 1) captured once as a HIP graph
    int              use_graph  = 1;
    bool             capturing  = false;
    hipGraph_t       graph      = nullptr;  // one steady-state step
    hipGraphExec_t   graph_exec = nullptr;
    hipGraph_t       graph_k      = nullptr;  // GRAPH_STEPS steady-state steps
    hipGraphExec_t   graph_exec_k = nullptr;
    int          np_domain = 0;
    CTime        jikan;
    int          step_path  = KH_PATH_FUSED;
    int          zeta_dirty = 1;  // zeta may hold non-kept modes
    int          zeta_lazy  = 0;  // zeta's non-kept modes are still to be formed from ureal (settle_zeta)
    // Lees-Edwards shear (kh_set_shear): DC rate or AC rate*cos(2 pi freq t);
    // deg = degree_oblique (strain of the oblique grid), sre = Shear_rate_eff
    int          shear = 0, shear_ac = 0;
    double       shear_rate = 0.0, shear_freq = 0.0, deg = 0.0, sre = 0.0, xyaspect = 1.0;
    double       strain = 0.0;                 // Shear_strain_realized
    double      *shear_pstress = nullptr;      // np x 6 stress partials of the sheared gather and LJ
    double      *shear_acc     = nullptr;      // Reynolds-stress partials (one per k_obl_to_cart workgroup)
    long long    shear_nacc    = 0;
    double      *shear_last    = nullptr;      // the last step's stress record (KH_SHEAR_RECORD doubles)
    double      *shear_hist    = nullptr;      // records of shear_hist_n steps from ts = shear_hist_base
    int          shear_hist_n = 0, shear_hist_base = 0;
    // rigid composite bodies (kh_set_rigid)
    int              rigid = 0;
    RigidArgs        ra{};
    int             *bid   = nullptr;   // np: body of each bead
    double          *flo   = nullptr;   // np x 3: LJ force from partners of higher index
    std::vector<int> sek_host;          // the packed Sekibun offsets (host copy, for the initial block)
    // measurement
    bool                               timing = false;
    std::map<std::string, KernelTimes> times;
    std::vector<PendingEvent>          pending;
    std::vector<hipEvent_t>            event_pool;
    std::vector<std::string>           names;
};

// A batch of runs of one grid (the serving loop's size classes): `capacity`
// equal arenas of `stride` bytes in one allocation, and one stream on which
// the joined runs step together. The batched launches skip slots whose
// device flag active[slot] is 0 (free slots, and runs still stepping alone).
// A batch step is captured once as a graph over every slot: which slots
// step is the device flags' business, and each joined run's particle
// arguments sit in a device table (SlotPart per slot), so runs join and
// leave without a new capture.
struct kh_batch {
    kh_params             P{};  // the class: grid, precision, dx, rho, eta, dt, FIX_CELL, a, a_xi
    int                   capacity = 0, np_cap = 0;
    size_t                stride   = 0;
    char                 *base     = nullptr;
    int                  *active   = nullptr;  // device flags, one per slot
    std::vector<int>      hactive;
    void                 *dslots = nullptr;    // device SlotPart<T> per slot
    std::vector<char>     hslots;              // its host copy
    hipStream_t           stream = nullptr;
    std::vector<kh_run *> slot_run;  // the run created in each slot (nullptr: free)
    bool                  dirty  = true;
    hipGraph_t            graph1 = nullptr, graphk = nullptr;
    hipGraphExec_t        exec1 = nullptr, execk = nullptr;
};

namespace {

// ---------------------------------------------------------------- timing
hipEvent_t take_event(kh_run *r) {
    if (!r->event_pool.empty()) {
        hipEvent_t e = r->event_pool.back();
        r->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int resolve_timing(kh_run *r) {
    if (r->pending.empty()) return KH_OK;
    KH_HIP(hipStreamSynchronize(r->stream));
    for (auto &p : r->pending) {
        float ms = 0.f;
        KH_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        KernelTimes &kt = r->times[p.name];
        kt.launches++;
        kt.ms += ms;
        r->event_pool.push_back(p.a);
        r->event_pool.push_back(p.b);
    }
    r->pending.clear();
    return KH_OK;
}

// Brackets one launch with events when timing is on.
struct Timed {
    kh_run     *r;
    const char *name;
    hipEvent_t  a = nullptr;
    Timed(kh_run *r_, const char *n) : r(r_), name(n) {
        if (r->timing && !r->capturing) {
            a = take_event(r);
            if (a) (void)hipEventRecord(a, r->stream);
        }
    }
    ~Timed() {
        if (r->timing && !r->capturing && a) {
            hipEvent_t b = take_event(r);
            if (b) {
                (void)hipEventRecord(b, r->stream);
                r->pending.push_back({name, a, b});
            }
        }
    }
};
#define KH_LAUNCH(name, call)        \
    do {                             \
        Timed t_(r, name);           \
        KH_HIP(call);                \
    } while (0)

// ------------------------------------------------------------ the step
// the run's occupancy bits of the 4-point z vectors
static inline Occ occ_of(const kh_run *r) { return Occ{(unsigned long long *)r->occ, r->occw}; }

template <typename T>
struct Step {
    static cpx<T> *S(kh_run *r, int i) { return (cpx<T> *)r->spec[i]; }
    static cpx<T> *Z(kh_run *r, int i) { return (cpx<T> *)r->zeta[i]; }
    static T      *U(kh_run *r, int i) { return (T *)r->ureal[i]; }
    static T      *UP(kh_run *r, int i) { return (T *)r->up[i]; }
    static T      *W(kh_run *r, int i) { return (T *)r->work[i]; }
    static T      *PS(kh_run *r) { return (T *)r->phisum; }

    static int r2c(kh_run *r, const T *in, cpx<T> *out) {
        KH_LAUNCH("fft_z_r2c", (fft_z_r2c<T>(r->g, in, out, r->stream)));
        KH_LAUNCH("fft_y_fwd", (fft_y<T, true>(r->g, out, r->stream)));
        KH_LAUNCH("fft_x_fwd", (fft_x<T, true>(r->g, out, r->stream)));
        return KH_OK;
    }
    static int c2r(kh_run *r, cpx<T> *in, T *out) {
        KH_LAUNCH("fft_x_inv", (fft_x<T, false>(r->g, in, r->stream)));
        KH_LAUNCH("fft_y_inv", (fft_y<T, false>(r->g, in, r->stream)));
        KH_LAUNCH("fft_z_c2r", (fft_z_c2r<T>(r->g, in, out, r->stream)));
        return KH_OK;
    }
    // The occupancy bits replace phi_sum as the mask of z_c2r and z_force
    // for plain spheres in FP64. In FP32 the bits cost the scatter more than
    // the mask reads they save (a 4-point vector is 16 bytes there, and
    // z_force's clearing stores into lines it has not read are slower).
    static bool occ_on(const kh_run *r) {
        return occ_kept<T>() && r->zclear && !r->slip && !r->rigid && !r->pp.obl;
    }
    static int zero(kh_run *r, T *p) {
        KH_LAUNCH("fill", (fill<T>(r->g.nreal, p, (T)0, r->stream)));
        return KH_OK;
    }

    // Zeta_k2u (operate_omega.h:134-137); in the step it is always preceded
    // by the 2/3 truncation of zeta (truncate = 1) and, after the NS solver,
    // by the FIX_CELL zeroing of uk_dc (fixmask)
    static int zeta_k2u(kh_run *r, int fixmask, int truncate = 1) {
        KH_LAUNCH("zeta_to_u", (zeta_to_u<T>(r->g, Z(r, 0), Z(r, 1), S(r, 0), S(r, 1), S(r, 2), r->ukdc, fixmask,
                                             truncate, r->stream)));
        for (int d = 0; d < 3; d++) KH_TRY(c2r(r, S(r, d), U(r, d)));
        return KH_OK;
    }

    // NS_solver_slavedEuler (fluid_solver.cxx:348-374) via Zeta_k2advection_k
    static int ns_solver(kh_run *r) {
        KH_TRY(zeta_k2u(r, 0));
        T *p[5] = {W(r, 0), W(r, 1), W(r, 2), W(r, 3), W(r, 4)};
        KH_LAUNCH("products", (products<T>(r->g.nreal, U(r, 0), U(r, 1), U(r, 2), p, r->stream)));
        for (int c = 0; c < 5; c++) KH_TRY(r2c(r, W(r, c), S(r, c)));
        cpx<T> *P5[5] = {S(r, 0), S(r, 1), S(r, 2), S(r, 3), S(r, 4)};
        KH_LAUNCH("adv_ns", (adv_ns<T>(r->g, P5, Z(r, 0), Z(r, 1), -r->nu * r->jikan.dt_fluid, 1. / r->nu,
                                       r->stream)));
        return KH_OK;
    }

    // U2zeta_k (operate_omega.h:208-211)
    static int u2zeta_k(kh_run *r) {
        for (int d = 0; d < 3; d++) KH_TRY(r2c(r, U(r, d), S(r, d)));
        KH_LAUNCH("u_to_zeta", (u_to_zeta<T>(r->g, S(r, 0), S(r, 1), S(r, 2), Z(r, 0), Z(r, 1), r->ukdc, r->stream)));
        return KH_OK;
    }

    static PCache<T> CP(kh_run *r) { return PCache<T>{(T *)r->cphi, r->cidx, r->ccode, r->ccnt}; }
    static int pos_update(kh_run *r, int ab2) {
        if (r->rigid)
            KH_LAUNCH("rigid_position", rigid_position(r->pp, r->dp, r->ra, ab2, r->jikan.dt_md, r->jikan.hdt_md,
                                                       r->stream));
        else
            KH_LAUNCH("position", position(r->pp, r->dp, ab2, r->jikan.dt_md, r->jikan.hdt_md, r->stream));
        r->pcache_valid = 0;
        return KH_OK;
    }
    // scatter phi (Make_phi_particle_sum) and fill the per-particle cache
    static int scatter_phi(kh_run *r) {
        KH_LAUNCH("phi_scatter", (phi_scatter<T>(r->pp, r->dp, PS(r), CP(r), 1, r->stream)));
        r->pcache_valid = 1;
        return KH_OK;
    }
    // the position integrator and the phi scatter in one launch
    static int move_scatter_phi(kh_run *r, int ab2) {
        KH_LAUNCH("phi_scatter", (phi_scatter<T>(r->pp, r->dp, PS(r), CP(r), 1, r->stream, 1, ab2,
                                                 r->jikan.dt_md, r->jikan.hdt_md)));
        r->pcache_valid = 1;
        return KH_OK;
    }
    static int ensure_cache(kh_run *r) {
        if (!r->pcache_valid) {
            KH_LAUNCH("phi_cache", (phi_scatter<T>(r->pp, r->dp, PS(r), CP(r), 0, r->stream)));
            r->pcache_valid = 1;
        }
        return KH_OK;
    }
    static int make_phi(kh_run *r) {
        KH_TRY(zero(r, PS(r)));
        return scatter_phi(r);
    }
    static int calc_f_hydro(kh_run *r) {
        KH_TRY(ensure_cache(r));
        T *u[3] = {U(r, 0), U(r, 1), U(r, 2)};
        KH_LAUNCH("f_hydro", (f_hydro<T>(r->pp, r->dp, u, PS(r), CP(r),
                                         (-r->dx3 * r->rho) / r->jikan.dt_fluid, r->stream)));
        return KH_OK;
    }
    // noise amplitudes of Add_random_force_thermostat (fluct.cxx:11-18)
    static void noise_amp(const kh_run *r, double &niv, double &nio) {
        const double zeta_drag = 6. * KH_PI * r->eta * r->radius;
        const double zeta_rot  = 8. * KH_PI * r->eta * r->radius * r->radius * r->radius;
        niv                    = kT_snap_v * sqrt(zeta_drag * r->jikan.dt_md * r->kbt * r->P.alpha_v);
        nio                    = kT_snap_o * sqrt(zeta_rot * r->jikan.dt_md * r->kbt * r->P.alpha_o);
    }
    // LJ + gravity, then the velocity update; with `noise` the thermostat's
    // kick follows in the same launch
    static int velocity_update(kh_run *r, int ab2, bool noise = false) {
        KH_LAUNCH("forces", forces(r->pp, r->dp, r->stream));
        if (r->rigid) {  // calc_Rigid_VOGs + set_Particle_Velocities (particle_solver.cxx:137-140, 274-277)
            KH_LAUNCH("rigid_velocity", rigid_velocity(r->pp, r->dp, r->ra, ab2, r->jikan.dt_md, r->jikan.hdt_md,
                                                       r->stream));
            return KH_OK;
        }
        double niv = 0.0, nio = 0.0;
        if (noise) noise_amp(r, niv, nio);
        KH_LAUNCH("velocity", velocity(r->pp, r->dp, ab2, ab2 ? r->jikan.hdt_md : r->jikan.dt_md, r->stream,
                                       noise ? 1 : 0, r->P.seed, r->P.stream_id, r->d_ts, niv, nio));
        return KH_OK;
    }
    static int thermo(kh_run *r) {
        double niv, nio;
        noise_amp(r, niv, nio);
        KH_LAUNCH("thermostat", thermostat(r->pp, r->dp, r->P.seed, r->P.stream_id, r->d_ts, niv, nio, r->stream));
        return KH_OK;
    }
    static int scatter_up(kh_run *r) {
        KH_TRY(ensure_cache(r));
        T *up3[3] = {UP(r, 0), UP(r, 1), UP(r, 2)};
        KH_LAUNCH("up_scatter", (up_scatter<T>(r->pp, r->dp, PS(r), CP(r), up3, r->stream)));
        return KH_OK;
    }
    static int make_up(kh_run *r) {
        for (int d = 0; d < 3; d++) KH_TRY(zero(r, UP(r, d)));
        return scatter_up(r);
    }
    // Make_f_particle_dt_sole (f_particle.h:60-67) into W[0..2]
    static int f_particle_sole(kh_run *r) {
        T *up3[3] = {UP(r, 0), UP(r, 1), UP(r, 2)}, *u[3] = {U(r, 0), U(r, 1), U(r, 2)};
        T *f[3]   = {W(r, 0), W(r, 1), W(r, 2)};
        KH_LAUNCH("f_nonsole", (f_nonsole<T>(r->g.nreal, up3, PS(r), u, f, r->stream)));
        for (int d = 0; d < 3; d++) KH_TRY(r2c(r, W(r, d), S(r, d)));
        KH_LAUNCH("solenoidal", (solenoidal<T>(r->g, S(r, 0), S(r, 1), S(r, 2), r->stream)));
        for (int d = 0; d < 3; d++) KH_TRY(c2r(r, S(r, d), W(r, d)));
        return KH_OK;
    }
    // The slip iteration of Time_evolution_hydro (sp_3d_ns.cxx:121-161):
    // Make_particle_momentum_factor, then until converged (or slip_iter_max)
    // slip scatter -> Solenoidal_u -> Calc_f_slip -> up += u -> (first pass:
    // LJ + gravity) -> velocity with the slip force; finally u := up. In the
    // fused step zeta also takes the curl of the projected slip field (its
    // u_hat is rebuilt from zeta later), and uk_dc its k = 0 mode.
    static int slip_iteration(kh_run *r, int ab2, bool fused) {
        T *u[3] = {U(r, 0), U(r, 1), U(r, 2)}, *up[3] = {UP(r, 0), UP(r, 1), UP(r, 2)};
        KH_LAUNCH("slip_moments", slip_moments<T>(r->pp, r->dp, u, r->stream));
        KH_LAUNCH("slip_update", slip_update(r->pp, r->dp, 0, 0, r->stream));
        const double coef = (r->dx3 * r->rho) / r->jikan.dt_fluid;
        const double h    = ab2 ? r->jikan.hdt_md : r->jikan.dt_md;
        int          iter = 0;
        for (;;) {
            for (int d = 0; d < 3; d++) KH_TRY(zero(r, up[d]));
            KH_LAUNCH("slip_scatter", slip_scatter<T>(r->pp, r->dp, u, up, r->stream));
            for (int d = 0; d < 3; d++) KH_TRY(r2c(r, up[d], S(r, d)));
            KH_LAUNCH("solenoidal", (solenoidal<T>(r->g, S(r, 0), S(r, 1), S(r, 2), r->stream)));
            if (fused)
                KH_LAUNCH("slip_zeta", (u_to_zeta<T>(r->g, S(r, 0), S(r, 1), S(r, 2), S(r, 3), S(r, 4), r->slipdc,
                                                     r->stream)));
            for (int d = 0; d < 3; d++) KH_TRY(c2r(r, S(r, d), up[d]));
            KH_LAUNCH("f_slip", f_slip<T>(r->pp, r->dp, up, coef, r->stream));
            KH_LAUNCH("add3", (add3<T>(r->g.nreal, up, u, r->stream)));
            if (iter == 0) KH_LAUNCH("forces", forces(r->pp, r->dp, r->stream));
            KH_LAUNCH("velocity_slip", velocity_slip(r->pp, r->dp, ab2, h, r->stream));
            iter++;
            KH_LAUNCH("slip_conv", slip_conv(r->pp, r->dp, r->slipeps, r->stream));
            double eps = 0.0;
            KH_HIP(hipMemcpyAsync(&eps, r->slipeps, sizeof(double), hipMemcpyDeviceToHost, r->stream));
            KH_HIP(hipStreamSynchronize(r->stream));
            if (eps < r->P.slip_tol || iter == r->P.slip_iter_max) {
                KH_LAUNCH("reset_forces", reset_forces(r->pp, r->dp, r->stream));
                break;
            }
            KH_LAUNCH("slip_update", slip_update(r->pp, r->dp, iter, 1, r->stream));
        }
        r->slip_iters = iter;
        r->slip_iters_host += iter;
        r->slip_loops_host++;
        for (int d = 0; d < 3; d++) std::swap(r->ureal[d], r->up[d]);  // Swap_mem(u, up)
        if (fused)
            KH_LAUNCH("zeta_add", (zeta_add<T>(r->g, Z(r, 0), Z(r, 1), S(r, 3), S(r, 4), r->ukdc, r->slipdc,
                                               r->stream)));
        return KH_OK;
    }

    // The same slip iteration in the fused step with its length decided on
    // the device: every iteration up to slip_iter_max is launched and those
    // after convergence return at once (slip_conv sets the flag the others
    // read), so squirmer steps are captured as graphs like the plain step.
    // The projection of the slip force field runs on the pruned transforms
    // (Solenoidal_u truncates to the 2/3 modes): forward z of the scattered
    // field (read at every point: the shell slip force is nonzero also where
    // phi has underflowed to 0), which leaves up zero as it reads it,
    // forward y, x_slip (projection, its curl and k = 0 mode, inverse x),
    // inverse y, inverse z into W only where phi_sum != 0. The projected
    // field is read only there: the slip force gather, and z_force, which
    // forms u = u* + W (u + up of the reference) where it needs it.
    static int slip_iteration_dev(kh_run *r, int ab2) {
        const Grid &g    = r->g;
        hipStream_t s    = r->stream;
        T          *u[3] = {U(r, 0), U(r, 1), U(r, 2)}, *up[3] = {UP(r, 0), UP(r, 1), UP(r, 2)};
        T          *w[3] = {W(r, 0), W(r, 1), W(r, 2)};
        cpx<T>     *s3[3] = {S(r, 0), S(r, 1), S(r, 2)};
        Comps<T>    c3{{S(r, 0), S(r, 1), S(r, 2), nullptr, nullptr}};
        ZIO<T>      zio{{S(r, 0), S(r, 1), S(r, 2)}, {w[0], w[1], w[2]}};
        const SlipLoop loop{r->slipflag, (unsigned long long *)((char *)r->slipflag + 16), r->P.slip_tol,
                            r->P.slip_iter_max};
        const int     *skip  = r->slipflag;
        const double   coef  = (r->dx3 * r->rho) / r->jikan.dt_fluid;
        const double   h     = ab2 ? r->jikan.hdt_md : r->jikan.dt_md;
        const T        scale = (T)(1.0 / ((double)g.nx * (double)g.ny * (double)g.nz));
        KH_LAUNCH("slip_moments", slip_moments<T>(r->pp, r->dp, u, s));
        KH_LAUNCH("slip_update", slip_update(r->pp, r->dp, 0, 0, s));
        KH_LAUNCH("slip_begin", slip_loop_begin(loop, s));
        for (int iter = 0; iter < r->P.slip_iter_max; iter++) {
            KH_LAUNCH("slip_scatter", slip_scatter<T>(r->pp, r->dp, u, up, s, skip));
            KH_LAUNCH("slip_z_fwd", (z_force<T>(g, up, nullptr, nullptr, s3, 1, nullptr, s, nullptr, skip)));
            KH_LAUNCH("slip_y_fwd", (cols_y<T, true>(g, c3, 3, 0, 1, s, skip)));
            KH_LAUNCH("slip_x", (x_slip<T>(g, s3, S(r, 3), S(r, 4), r->slipdc, skip, s)));
            KH_LAUNCH("slip_y_inv", (cols_y<T, false>(g, c3, 3, 1, 0, s, skip)));
            KH_LAUNCH("slip_z_inv", (zw_c2r<T>(g.nz, zio, 3, (long long)g.nx * g.ny, scale, g.trn_z, g.nzp, s, PS(r),
                                               skip)));
            KH_LAUNCH("f_slip", f_slip<T>(r->pp, r->dp, w, coef, s, skip));
            if (iter == 0) KH_LAUNCH("forces", forces(r->pp, r->dp, s));
            KH_LAUNCH("velocity_slip", velocity_slip(r->pp, r->dp, ab2, h, s, skip));
            KH_LAUNCH("slip_conv", slip_conv(r->pp, r->dp, r->slipeps, s, &loop));
            if (iter + 1 < r->P.slip_iter_max)
                KH_LAUNCH("slip_update", slip_update(r->pp, r->dp, iter + 1, 1, s, skip));
        }
        KH_LAUNCH("reset_forces", reset_forces(r->pp, r->dp, s));
        KH_LAUNCH("zeta_add", (zeta_add<T>(g, Z(r, 0), Z(r, 1), S(r, 3), S(r, 4), r->ukdc, r->slipdc, s, 1)));
        return KH_OK;
    }

    static int add_f(kh_run *r) {
        T *u[3] = {U(r, 0), U(r, 1), U(r, 2)}, *f[3] = {W(r, 0), W(r, 1), W(r, 2)};
        KH_LAUNCH("add3", (add3<T>(r->g.nreal, u, f, r->stream)));
        return KH_OK;
    }

    // Time_evolution_hydro (sp_3d_ns.cxx:41-194) as the fused, pruned
    // pipeline of kh_fused.hip. Differences from step_reference_order are
    // round-off only: FFT(IFFT(a)) = a for the band-limited fields, so the
    // final U2zeta_k of u* + f is taken directly from u*_hat + P f_hat.
    // one step + the device step counter
    static int step(kh_run *r) {
        KH_TRY(r->shear ? step_shear(r) : step_body(r));
        // the fused body ticks inside z_force (or right after it)
        if (r->shear || r->step_path == KH_PATH_REFERENCE_ORDER) KH_LAUNCH("tick", tick(r->d_ts, r->stream));
        if (r->shear) {
            // Mean_shear_stress after the step (sp_3d_ns.cxx:962-964, sp_3d_ns.h:247-330)
            const double dt = r->jikan.dt_fluid, vol = r->pp.L[0] * r->pp.L[1] * r->pp.L[2];
            r->strain += r->sre * dt;
            const ShearRec h{r->jikan.time + dt, r->sre, r->deg, r->strain, dt, r->eta, 1.0 / vol, r->dx3 / vol};
            const int slot = r->jikan.ts - r->shear_hist_base;
            double   *hist = (r->shear_hist && slot >= 0 && slot < r->shear_hist_n) ? r->shear_hist + 12 * slot : nullptr;
            KH_LAUNCH("obl_record", obl_record(r->pp, r->shear_acc, r->shear_nacc, h, r->shear_last, hist, r->stream));
        }
        return records(r);
    }
    // the per-step trajectory record and Output_udf frames, when asked for
    static int records(kh_run *r) {
        if (r->traj)
            KH_LAUNCH("record", record(r->pp, r->dp, r->d_ts, r->traj_base, r->traj_steps, r->traj, r->stream));
        if (r->frames)
            KH_LAUNCH("frame", frame(r->pp, r->dp, r->d_ts, r->frame_ts0, r->frame_every, r->frame_n, r->frames,
                                     r->stream));
        return KH_OK;
    }

    // Reset_phi / Reset_u (the previous fused step's z_force already left
    // u_p and phi_sum zero), then the position update and the phi scatter
    static int particles_begin(kh_run *r, int ab2) {
        if (!(r->zclear && r->pclean)) {
            for (int d = 0; d < 3; d++) KH_TRY(zero(r, UP(r, d)));
            KH_TRY(zero(r, PS(r)));
            KH_LAUNCH("fill", (fill<double>((long long)r->occw * r->g.nx * r->g.ny, (double *)r->occ, 0.0, r->stream)));
        }
        r->pclean = false;
        if (r->rigid) {  // the bodies move, then their beads scatter phi
            KH_TRY(pos_update(r, ab2));
            return scatter_phi(r);
        }
        return move_scatter_phi(r, ab2);
    }

    // NS_solver_slavedEuler, FIX_CELL, truncation and Zeta_k2u of the fused
    // step, up to u* in real space (written only where mask != 0 when a mask
    // is given), for the run r or, when g carries a batch, for every run of
    // the batch from r on: one launch per pass covers them all.
    static int fused_fluid(kh_run *r, const Grid &g, const T *mask, Occ occ = {}) {
        hipStream_t s     = r->stream;
        cpx<T>     *s3[3] = {S(r, 0), S(r, 1), S(r, 2)};
        Comps<T>    c5{{S(r, 0), S(r, 1), S(r, 2), S(r, 3), S(r, 4)}};
        // buffer of each advection product (u2u3, u3u1, u1u2, u2^2-u3^2,
        // u3^2-u1^2): z_adv reads u1, u2, u3 from S0, S1, S2 and writes
        // each product only into a buffer it has finished reading
        cpx<T>  *q5[5] = {S(r, 1), S(r, 2), S(r, 0), S(r, 4), S(r, 3)};
        Comps<T> cq5{{q5[0], q5[1], q5[2], q5[3], q5[4]}};
        // NS_solver_slavedEuler: zeta -> u -> products -> -Omega -> slaved Euler
        KH_LAUNCH("x_zeta2u_inv", (x_zeta2u_inv<T>(g, Z(r, 0), Z(r, 1), s3, r->ukdc, 0, s)));
        KH_LAUNCH("y_inv", (cols_y<T, false>(g, c5, 3, 1, 0, s)));
        KH_LAUNCH("z_adv", (z_adv<T>(g, s3, q5, s)));
        KH_LAUNCH("y_fwd_adv", (cols_y<T, true>(g, cq5, 5, 0, 1, s)));
        if constexpr (sizeof(T) == 8) {  // forward x + slaved Euler in one pass
            KH_LAUNCH("x_adv_ns", (x_adv_ns<T>(g, cq5, Z(r, 0), Z(r, 1), -r->nu * r->jikan.dt_fluid, 1. / r->nu, s)));
        } else {  // FP32: the split passes keep more workgroups resident
            KH_LAUNCH("x_fwd_adv", (cols_x<T, true>(g, cq5, 5, 0, 1, s)));
            KH_LAUNCH("adv_ns", (adv_ns<T>(g, q5, Z(r, 0), Z(r, 1), -r->nu * r->jikan.dt_fluid, 1. / r->nu, s)));
        }
        // FIX_CELL, truncation, Zeta_k2u -> u* in real space
        KH_LAUNCH("x_zeta2u_inv", (x_zeta2u_inv<T>(g, Z(r, 0), Z(r, 1), s3, r->ukdc, r->fixmask, s)));
        KH_LAUNCH("y_inv", (cols_y<T, false>(g, c5, 3, 1, 0, s)));
        ZIO<T> zio{{S(r, 0), S(r, 1), S(r, 2)}, {U(r, 0), U(r, 1), U(r, 2)}};
        KH_LAUNCH("z_c2r", (fft_z_c2r_multi<T>(g, zio, 3, s, g.trn_z, mask, occ)));
        return KH_OK;
    }
    // f = P(up - phi u*), u = u* + f, zeta = curl u (u* = u + us2 when us2
    // is given; up_is_f: the particle step left f itself in up), and the
    // device step counter; for r or r's batch as above
    static int fused_force(kh_run *r, const Grid &g, T *const us2[3], bool up_is_f, Occ occ = {}) {
        hipStream_t s      = r->stream;
        cpx<T>     *s3[3]  = {S(r, 0), S(r, 1), S(r, 2)};
        Comps<T>    c5{{S(r, 0), S(r, 1), S(r, 2), S(r, 3), S(r, 4)}};
        T          *up3[3] = {UP(r, 0), UP(r, 1), UP(r, 2)}, *u3[3] = {U(r, 0), U(r, 1), U(r, 2)};
        KH_LAUNCH("z_force", (z_force<T>(g, up3, PS(r), up_is_f ? nullptr : u3, s3, r->zclear ? 3 : 0,
                                         r->zclear ? r->d_ts : nullptr, s, us2, nullptr, occ)));
        if (!r->zclear) KH_LAUNCH("tick", tick(r->d_ts, s, &g));
        KH_LAUNCH("y_fwd_force", (cols_y<T, true>(g, c5, 3, 0, 1, s)));
        KH_LAUNCH("x_f_zeta", (x_f_zeta<T>(g, s3, Z(r, 0), Z(r, 1), r->ukdc, s)));
        return KH_OK;
    }
    // F_H, LJ + gravity, velocity (+ noise), u_p scatter of a run of plain
    // spheres: one launch
    static int plain_particles(kh_run *r, int ab2) {
        T     *u[3] = {U(r, 0), U(r, 1), U(r, 2)}, *up[3] = {UP(r, 0), UP(r, 1), UP(r, 2)};
        double niv = 0.0, nio = 0.0;
        if (r->kbt > 0.) noise_amp(r, niv, nio);
        KH_LAUNCH("particle_step",
                  (particle_step<T>(r->pp, r->dp, u, PS(r), CP(r), (-r->dx3 * r->rho) / r->jikan.dt_fluid, ab2,
                                    ab2 ? r->jikan.hdt_md : r->jikan.dt_md, r->kbt > 0. ? 1 : 0, r->P.seed,
                                    r->P.stream_id, r->d_ts, niv, nio, up, r->stream)));
        return KH_OK;
    }

    static int step_body(kh_run *r) {
        if (r->step_path == KH_PATH_REFERENCE_ORDER) return step_reference_order(r);
        const int   ab2 = r->jikan.ts != 0;
        const Grid &g   = r->g;
        hipStream_t s   = r->stream;
        if (r->zeta_dirty) {
            KH_LAUNCH("zeta_clean", (zeta_clean<T>(g, Z(r, 0), Z(r, 1), s)));
            r->zeta_dirty = 0;
        }
        // The particles move and scatter phi first (the position update
        // reads only the particle state, so it commutes with the fluid
        // part): phi_sum is then known when u* is formed, and u* is written
        // only where it is read. Squirmer runs with the slip loop decided on
        // the host keep the reference's order.
        const bool early = !r->slip || r->slip_dev;
        // plain spheres in FP64: the phi scatter marks the 4-point vectors it
        // reaches, and z_c2r and z_force read those bits instead of phi_sum
        const Occ occ = occ_on(r) ? occ_of(r) : Occ{};
        r->pp.occ     = occ;
        if (early) KH_TRY(particles_begin(r, ab2));
        // squirmers read u* on the whole slip shell (|r - a| < xi/2), which
        // reaches points where phi has underflowed to 0: u* is written in full
        KH_TRY(fused_fluid(r, g, early && !r->slip ? PS(r) : nullptr, occ));
        if (!early) KH_TRY(particles_begin(r, ab2));
        if (r->slip_dev) {  // the slip loop leaves up zero
            KH_TRY(calc_f_hydro(r));
            KH_TRY(slip_iteration_dev(r, ab2));
            if (r->kbt > 0.) KH_TRY(thermo(r));
            KH_TRY(scatter_up(r));
        } else if (r->slip) {
            KH_TRY(calc_f_hydro(r));
            KH_TRY(slip_iteration(r, ab2, true));
            for (int d = 0; d < 3; d++) KH_TRY(zero(r, UP(r, d)));
            if (r->kbt > 0.) KH_TRY(thermo(r));
            KH_TRY(scatter_up(r));
        } else if (r->rigid) {  // bead F_H, LJ, the bodies' update, u_p from the beads' rigid motion
            KH_TRY(calc_f_hydro(r));
            KH_TRY(velocity_update(r, ab2));
            KH_TRY(scatter_up(r));
        } else {
            KH_TRY(plain_particles(r, ab2));
        }
        T *w3[3] = {W(r, 0), W(r, 1), W(r, 2)};  // squirmers: u = u* + the projected slip field
        KH_TRY(fused_force(r, g, r->slip_dev ? w3 : nullptr, !r->slip && !r->rigid, occ));
        r->pclean = r->zclear;
        return KH_OK;
    }

    // One steady-state step of every joined run of a batch (plain spheres
    // on one grid, ts >= 1): one launch per pass over all slots. v is a view
    // of slot 0 (a run's state with its device pointers moved to slot 0's
    // arena) that supplies the fluid passes' run-0 buffers and the class's
    // constants; the particle launches take each slot's own arguments from
    // the batch's SlotPart table.
    static int step_batch(kh_batch *b, kh_run *v) {
        kh_run     *r = v;
        Grid        g = r->g;
        g.nbatch      = b->capacity;
        g.bstride     = (long long)b->stride;
        g.bactive     = b->active;
        hipStream_t s = r->stream;
        const SlotPart<T> *sp = (const SlotPart<T> *)b->dslots;
        if (!r->zclear) {  // Reset_phi / Reset_u
            for (int d = 0; d < 3; d++) KH_LAUNCH("fill", (fill<T>(g.nreal, UP(r, d), (T)0, s, &g)));
            KH_LAUNCH("fill", (fill<T>(g.nreal, PS(r), (T)0, s, &g)));
        }
        KH_LAUNCH("phi_scatter", (phi_scatter_batch<T>(sp, b->active, b->capacity, b->np_cap, r->jikan.dt_md,
                                                       r->jikan.hdt_md, s)));
        const Occ occ = occ_on(r) ? occ_of(r) : Occ{};
        KH_TRY(fused_fluid(r, g, PS(r), occ));
        KH_LAUNCH("particle_step", (particle_step_batch<T>(sp, b->active, b->capacity, b->np_cap,
                                                           (-r->dx3 * r->rho) / r->jikan.dt_fluid, r->jikan.hdt_md,
                                                           s)));
        KH_TRY(fused_force(r, g, nullptr, true, occ));
        KH_LAUNCH("records", (records_batch<T>(sp, b->active, b->capacity, s)));
        return KH_OK;
    }
    // a run's arguments of the batched particle launches
    static SlotPart<T> slot_part(kh_run *r) {
        SlotPart<T> a{};
        a.pp     = r->pp;
        a.pp.occ = occ_on(r) ? occ_of(r) : Occ{};
        a.P      = r->dp;
        a.phisum = PS(r);
        a.pc     = CP(r);
        for (int d = 0; d < 3; d++) {
            a.u[d]  = U(r, d);
            a.up[d] = UP(r, d);
        }
        double niv = 0.0, nio = 0.0;
        if (r->kbt > 0.) noise_amp(r, niv, nio);
        a.nz          = Noise{r->kbt > 0. ? 1 : 0, r->P.seed, r->P.stream_id, r->d_ts, niv, nio};
        a.traj        = r->traj;
        a.traj_base   = r->traj_base;
        a.traj_steps  = r->traj_steps;
        a.frames      = r->frames;
        a.frame_ts0   = r->frame_ts0;
        a.frame_every = r->frame_every;
        a.frame_n     = r->frame_n;
        return a;
    }

    // Time_evolution_hydro in the reference's own operation order, one
    // kernel sequence per reference function.
    static int step_reference_order(kh_run *r) {
        const int ab2 = r->jikan.ts != 0;
        r->pclean     = false;
        KH_TRY(ns_solver(r));
        KH_TRY(zeta_k2u(r, r->fixmask));  // FIX_CELL, truncation, u*
        KH_TRY(pos_update(r, ab2));
        KH_TRY(make_phi(r));
        KH_TRY(calc_f_hydro(r));
        if (r->slip)
            KH_TRY(slip_iteration(r, ab2, false));
        else
            KH_TRY(velocity_update(r, ab2));
        if (r->kbt > 0.) KH_TRY(thermo(r));
        KH_TRY(make_up(r));
        KH_TRY(f_particle_sole(r));
        KH_TRY(add_f(r));
        KH_TRY(u2zeta_k(r));
        return KH_OK;
    }

    // Time_evolution_hydro_OBL (sp_3d_ns.cxx:328-433) with the oblique-grid
    // operators of kh_shear.hip. The strain and the shear rate are host
    // state passed to each launch by value; this step's values are known
    // before it starts, so the particles move and scatter phi first (the
    // position update reads only the particle state: it commutes with the
    // fluid part), and the Cartesian u* and f = u_p - phi u* are then formed
    // only where phi_sum is nonzero. Shear_rate_eff is the set rate:
    // Calc_shear_rate_eff (sp_3d_ns.h:189-203) adds the mean of a centred
    // difference over the periodic grid, which telescopes to zero.
    // Inverse transform of three 2/3-truncated spectra: the x pass on the
    // kept (ky, kz) columns reading kept kx rows, the y pass on kz < TRN_Z
    // reading kept ky rows, the z pass reading kz < TRN_Z.
    static int c2r_kept3(kh_run *r, cpx<T> *const sp[3], T *const out[3]) {
        const Grid &g = r->g;
        Comps<T>    c{{sp[0], sp[1], sp[2], nullptr, nullptr}};
        KH_LAUNCH("x_inv_kept", (cols_x<T, false>(g, c, 3, 1, 0, r->stream)));
        KH_LAUNCH("y_inv_kept", (cols_y<T, false>(g, c, 3, 1, 0, r->stream)));
        ZIO<T> zio{{sp[0], sp[1], sp[2]}, {out[0], out[1], out[2]}};
        KH_LAUNCH("z_c2r", (fft_z_c2r_multi<T>(g, zio, 3, r->stream, g.trn_z)));
        return KH_OK;
    }
    // The same inverse transform of omega's three spectra, then u x omega
    // (contravariant u and omega, covariant product) into out: fused into
    // the z pass where the wavefront engine has the length, else a separate
    // pass. u holds the real-space u.
    static int c2r_kept3_cross(kh_run *r, cpx<T> *const sp[3], T *const u[3], T *const out[3]) {
        const Grid &g = r->g;
        if (!zw_cross_supported(g.nz)) {
            KH_TRY(c2r_kept3(r, sp, out));
            KH_LAUNCH("obl_cross", (obl_cross<T>(g.nreal, u, out, r->stream)));
            return KH_OK;
        }
        Comps<T> c{{sp[0], sp[1], sp[2], nullptr, nullptr}};
        KH_LAUNCH("x_inv_kept", (cols_x<T, false>(g, c, 3, 1, 0, r->stream)));
        KH_LAUNCH("y_inv_kept", (cols_y<T, false>(g, c, 3, 1, 0, r->stream)));
        const T scale = (T)(1.0 / ((double)g.nx * (double)g.ny * (double)g.nz));
        ZIO<T>  zio{{sp[0], sp[1], sp[2]}, {out[0], out[1], out[2]}};
        KH_LAUNCH("z_c2r_cross",
                  (zw_c2r_cross<T>(g.nz, zio, u, (long long)g.nx * g.ny, scale, g.trn_z, g.nzp, r->stream)));
        return KH_OK;
    }
    // Forward transform of three fields whose consumer reads kept modes only:
    // y and x passes write only the kept rows of the kept columns.
    static int r2c_kept3(kh_run *r, T *const in[3], cpx<T> *const sp[3]) {
        const Grid &g = r->g;
        for (int d = 0; d < 3; d++) KH_LAUNCH("fft_z_r2c", (fft_z_r2c<T>(g, in[d], sp[d], r->stream, g.trn_z)));
        Comps<T> c{{sp[0], sp[1], sp[2], nullptr, nullptr}};
        KH_LAUNCH("y_fwd_kept", (cols_y<T, true>(g, c, 3, 0, 1, r->stream)));
        KH_LAUNCH("x_fwd_kept", (cols_x<T, true>(g, c, 3, 0, 1, r->stream)));
        return KH_OK;
    }

    static int step_shear(kh_run *r) {
        const int    ab2 = r->jikan.ts != 0;
        const Grid  &g   = r->g;
        hipStream_t  s   = r->stream;
        const double gd  = r->deg, dt = r->jikan.dt_fluid;
        cpx<T>      *s3[3] = {S(r, 0), S(r, 1), S(r, 2)};
        T           *u3[3] = {U(r, 0), U(r, 1), U(r, 2)}, *w3[3] = {W(r, 0), W(r, 1), W(r, 2)};
        r->pclean          = false;
        // this step's rate and strain; past +-NX/NY the grid is remapped
        // (Reset_U_OBL) after the fluid part
        const double sre = r->shear_ac ? r->shear_rate * cos(2. * KH_PI * r->shear_freq * dt * r->jikan.ts)
                                       : r->shear_rate;
        double       deg   = gd + sre * dt;
        const int    remap = deg >= r->xyaspect ? 1 : (deg <= -r->xyaspect ? -1 : 0);
        deg -= remap * r->xyaspect;
        r->pp.obl_g    = deg;
        r->pp.obl_jump = sre * r->pp.L[1];
        KH_TRY(pos_update(r, ab2));
        KH_TRY(make_phi(r));
        // NS_solver_slavedEuler_Shear_OBL at the previous strain and rate:
        // u and omega of the truncated zeta, u x omega, the shear terms,
        // curl, slaved Euler
        // (zeta is truncated here, so u, omega and the kept modes of u x omega
        // take the pruned transforms)
        // (zeta's modes beyond the 2/3 rule are left as they are: nothing in
        // the step reads them before U_k2zeta_k_OBL rewrites every mode)
        KH_LAUNCH("obl_zeta_to_u", (obl_zeta_to_u<T>(g, Z(r, 0), Z(r, 1), s3[0], s3[1], s3[2], S(r, 3), r->ukdc, 0, 0,
                                                     gd, s)));
        KH_TRY(c2r_kept3(r, s3, u3));
        KH_LAUNCH("obl_zeta_to_omega", (obl_zeta_to_omega<T>(g, Z(r, 0), Z(r, 1), s3, s)));
        KH_TRY(c2r_kept3_cross(r, s3, u3, w3));
        KH_TRY(r2c_kept3(r, w3, s3));
        KH_LAUNCH("obl_adv_ns", (obl_adv_ns<T>(g, s3, S(r, 3), Z(r, 0), Z(r, 1), r->sre, gd, -r->nu * dt, 1. / r->nu,
                                               s)));
        // FIX_CELL, Zeta_k2u_k_OBL, U_k2u: u* on the oblique grid
        KH_LAUNCH("obl_zeta_to_u", (obl_zeta_to_u<T>(g, Z(r, 0), Z(r, 1), s3[0], s3[1], s3[2], nullptr, r->ukdc,
                                                     r->fixmask, 0, gd, s)));
        KH_TRY(c2r_kept3(r, s3, u3));
        if (remap) {
            KH_LAUNCH("obl_remap", (obl_remap<T>(g, u3, w3, remap * r->xyaspect, s)));
            for (int d = 0; d < 3; d++) std::swap(r->ureal[d], r->work[d]);
            for (int d = 0; d < 3; d++) {
                u3[d] = U(r, d);
                w3[d] = W(r, d);
            }
        }
        r->sre = sre;
        r->deg = deg;
        // u* on the Cartesian grid with the mean flow (ucp) into W, where phi_sum != 0
        KH_LAUNCH("obl_to_cart", (obl_to_cart<T>(g, u3, w3, deg, sre, PS(r), r->shear_acc, r->P.dx, s)));
        // F_H on ucp, velocities, noise, u_p
        KH_LAUNCH("f_hydro", (f_hydro<T>(r->pp, r->dp, w3, PS(r), CP(r), (-r->dx3 * r->rho) / dt, s)));
        KH_TRY(velocity_update(r, ab2));
        if (r->kbt > 0.) KH_TRY(thermo(r));
        KH_TRY(make_up(r));
        // f = u_p - phi ucp onto the oblique grid, u += f; project; zeta
        T *up3[3] = {UP(r, 0), UP(r, 1), UP(r, 2)};
        KH_LAUNCH("obl_add_f", (obl_add_f<T>(g, up3, PS(r), w3, u3, deg, s)));
        // the kept modes only: the next sheared step reads no others, and
        // settle() forms the rest from u (left in ureal) when anything else
        // is about to read or change them
        KH_TRY(r2c_kept3(r, u3, s3));
        KH_LAUNCH("obl_u_to_zeta", (obl_u_to_zeta<T>(g, s3[0], s3[1], s3[2], Z(r, 0), Z(r, 1), r->ukdc, deg, s, 1)));
        r->zeta_dirty = 1;  // zeta keeps the modes of f beyond the 2/3 rule
        r->zeta_lazy  = 1;
        return KH_OK;
    }

    // The rest of the sheared step's U_k2u + Solenoidal_uk_OBL +
    // U_k2zeta_k_OBL: every mode of zeta from the u the step left in ureal,
    // at the step's strain (the kept modes come out as the step formed them).
    static int settle(kh_run *r) {
        const Grid &g     = r->g;
        cpx<T>     *s3[3] = {S(r, 0), S(r, 1), S(r, 2)};
        for (int d = 0; d < 3; d++) KH_TRY(r2c(r, U(r, d), s3[d]));
        KH_LAUNCH("obl_u_to_zeta",
                  (obl_u_to_zeta<T>(g, s3[0], s3[1], s3[2], Z(r, 0), Z(r, 1), r->ukdc, r->deg, r->stream)));
        r->zeta_lazy = 0;
        return KH_OK;
    }

    // Init_zeta_k on a fluid at rest + sp_3d_ns.cxx:809-826 (the plain
    // functions, whatever the run: the reference calls them for sheared runs too)
    static int init_fluid(kh_run *r) {
        const int obl = r->pp.obl;
        r->pp.obl     = 0;
        const int st  = init_fluid_plain(r);
        r->pp.obl     = obl;
        return st;
    }
    static int init_fluid_plain(kh_run *r) {
        r->pclean = false;
        KH_HIP(hipMemsetAsync(r->zeta[0], 0, sizeof(cpx<T>) * r->g.nspec, r->stream));
        KH_HIP(hipMemsetAsync(r->zeta[1], 0, sizeof(cpx<T>) * r->g.nspec, r->stream));
        KH_HIP(hipMemsetAsync(r->ukdc, 0, sizeof(double) * 3, r->stream));
        KH_TRY(make_phi(r));
        KH_TRY(make_up(r));
        KH_TRY(zeta_k2u(r, 0));
        KH_TRY(f_particle_sole(r));
        KH_TRY(add_f(r));
        KH_TRY(u2zeta_k(r));
        KH_HIP(hipMemsetAsync(r->ukdc, 0, sizeof(double) * 3, r->stream));
        return KH_OK;
    }

    // single functions of the sheared step on the run's state (strain and
    // rate as kh_set_shear_state left them)
    static int apply_obl(kh_run *r, int fn) {
        const Grid &g      = r->g;
        hipStream_t s      = r->stream;
        cpx<T>     *s3[3]  = {S(r, 0), S(r, 1), S(r, 2)};
        T          *u3[3]  = {U(r, 0), U(r, 1), U(r, 2)}, *w3[3] = {W(r, 0), W(r, 1), W(r, 2)};
        const double dt    = r->jikan.dt_fluid;
        switch (fn) {
            case KH_FN_OBL_NS_SOLVER:  // Zeta_k2advection_k_OBL + slaved Euler, as in the step
                KH_LAUNCH("obl_zeta_to_u", (obl_zeta_to_u<T>(g, Z(r, 0), Z(r, 1), s3[0], s3[1], s3[2], S(r, 3),
                                                             r->ukdc, 0, 1, r->deg, s)));
                KH_TRY(c2r_kept3(r, s3, u3));
                KH_LAUNCH("obl_zeta_to_omega", (obl_zeta_to_omega<T>(g, Z(r, 0), Z(r, 1), s3, s)));
                KH_TRY(c2r_kept3_cross(r, s3, u3, w3));
                KH_TRY(r2c_kept3(r, w3, s3));
                KH_LAUNCH("obl_adv_ns", (obl_adv_ns<T>(g, s3, S(r, 3), Z(r, 0), Z(r, 1), r->sre, r->deg,
                                                       -r->nu * dt, 1. / r->nu, s)));
                return KH_OK;
            case KH_FN_OBL_ZETA_K2U:  // Zeta_k2u_k_OBL + U_k2u on every mode of zeta
                KH_LAUNCH("obl_zeta_to_u", (obl_zeta_to_u<T>(g, Z(r, 0), Z(r, 1), s3[0], s3[1], s3[2], nullptr,
                                                             r->ukdc, 0, -1, r->deg, s)));
                for (int d = 0; d < 3; d++) KH_TRY(c2r(r, s3[d], u3[d]));
                return KH_OK;
            case KH_FN_OBL_U2ZETA_K:  // U2u_k + Solenoidal_uk_OBL + U_k2zeta_k_OBL
                for (int d = 0; d < 3; d++) KH_TRY(r2c(r, u3[d], s3[d]));
                KH_LAUNCH("obl_u_to_zeta", (obl_u_to_zeta<T>(g, s3[0], s3[1], s3[2], Z(r, 0), Z(r, 1), r->ukdc,
                                                             r->deg, s)));
                return KH_OK;
            case KH_FN_OBL_OBLIQUE2U:  // U_oblique2u on u, in place
                KH_LAUNCH("obl_to_cart", (obl_to_cart<T>(g, u3, w3, r->deg, r->sre, nullptr, nullptr, r->P.dx, s)));
                for (int d = 0; d < 3; d++) std::swap(r->ureal[d], r->work[d]);
                return KH_OK;
        }
        return fail(KH_ERR_ARG, "kh_apply: unknown function %d", fn);
    }

    static int apply(kh_run *r, int fn) {
        r->zeta_dirty = 1;
        r->pclean     = false;
        switch (fn) {
            case KH_FN_NS_SOLVER: return ns_solver(r);
            case KH_FN_ZETA_K2U: return zeta_k2u(r, 0, 0);
            case KH_FN_U2ZETA_K: return u2zeta_k(r);
            case KH_FN_POSITION_EULER: return pos_update(r, 0);
            case KH_FN_POSITION_AB2: return pos_update(r, 1);
            case KH_FN_MAKE_PHI: return make_phi(r);
            case KH_FN_F_HYDRO: return calc_f_hydro(r);
            case KH_FN_VELOCITY_EULER: return velocity_update(r, 0);
            case KH_FN_VELOCITY_AB2: return velocity_update(r, 1);
            case KH_FN_THERMOSTAT: return thermo(r);
            case KH_FN_MAKE_UP: return make_up(r);
            case KH_FN_F_PARTICLE_SOLE: return f_particle_sole(r);
            case KH_FN_ADD_F_PARTICLE: return add_f(r);
            case KH_FN_SLIP_MOMENTUM: {
                T *u[3] = {U(r, 0), U(r, 1), U(r, 2)};
                KH_LAUNCH("slip_moments", slip_moments<T>(r->pp, r->dp, u, r->stream));
                return KH_OK;
            }
            case KH_FN_SLIP_SCATTER: {
                T *u[3] = {U(r, 0), U(r, 1), U(r, 2)}, *up[3] = {UP(r, 0), UP(r, 1), UP(r, 2)};
                for (int d = 0; d < 3; d++) KH_TRY(zero(r, up[d]));
                KH_LAUNCH("slip_scatter", slip_scatter<T>(r->pp, r->dp, u, up, r->stream));
                return KH_OK;
            }
            case KH_FN_SLIP_ITERATION:
                if (!r->slip) return fail(KH_ERR_STATE, "no slip species in this run");
                return slip_iteration(r, r->jikan.ts != 0, false);
            case KH_FN_OBL_NS_SOLVER:
            case KH_FN_OBL_ZETA_K2U:
            case KH_FN_OBL_U2ZETA_K:
            case KH_FN_OBL_OBLIQUE2U:
                if (!r->shear) return fail(KH_ERR_STATE, "not a sheared run (kh_set_shear)");
                return apply_obl(r, fn);
            default: return fail(KH_ERR_ARG, "kh_apply: unknown function %d", fn);
        }
    }
};

// ------------------------------------------------------------ transfers
// host reference layout [NX][NY][NZ_] (real in k < NZ) <-> device [NX][NY][NZ]
int upload_real(kh_run *r, void *dev, const double *host) {
    const Grid &g   = r->g;
    const int   nz_ = 2 * g.nzh;
    if (!r->mixed) {
        KH_HIP(hipMemcpy2DAsync(dev, sizeof(double) * g.nz, host, sizeof(double) * nz_, sizeof(double) * g.nz,
                                (size_t)g.nx * g.ny, hipMemcpyHostToDevice, r->stream));
        KH_HIP(hipStreamSynchronize(r->stream));
    } else {
        std::vector<float> tmp((size_t)g.nreal);
        for (long long row = 0; row < (long long)g.nx * g.ny; row++)
            for (int k = 0; k < g.nz; k++) tmp[(size_t)row * g.nz + k] = (float)host[(size_t)row * nz_ + k];
        KH_HIP(hipMemcpyAsync(dev, tmp.data(), sizeof(float) * tmp.size(), hipMemcpyHostToDevice, r->stream));
        KH_HIP(hipStreamSynchronize(r->stream));
    }
    return KH_OK;
}
int download_real(kh_run *r, double *host, const void *dev) {
    const Grid &g   = r->g;
    const int   nz_ = 2 * g.nzh;
    if (!r->mixed) {
        KH_HIP(hipMemcpy2DAsync(host, sizeof(double) * nz_, dev, sizeof(double) * g.nz, sizeof(double) * g.nz,
                                (size_t)g.nx * g.ny, hipMemcpyDeviceToHost, r->stream));
        KH_HIP(hipStreamSynchronize(r->stream));
    } else {
        std::vector<float> tmp((size_t)g.nreal);
        KH_HIP(hipMemcpyAsync(tmp.data(), dev, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost, r->stream));
        KH_HIP(hipStreamSynchronize(r->stream));
        for (long long row = 0; row < (long long)g.nx * g.ny; row++)
            for (int k = 0; k < g.nz; k++) host[(size_t)row * nz_ + k] = tmp[(size_t)row * g.nz + k];
    }
    for (long long row = 0; row < (long long)g.nx * g.ny; row++)
        for (int k = g.nz; k < nz_; k++) host[(size_t)row * nz_ + k] = 0.0;
    return KH_OK;
}
// spectral: host rows of NZ_ = 2 nzh doubles ((re,im) of kz at [2kz]),
// device rows nzp complex apart (padding left untouched); FP32 runs convert
int upload_spec(kh_run *r, void *dev, const double *host) {
    const Grid     &g    = r->g;
    const long long rows = (long long)g.nx * g.ny;
    if (!r->mixed) {
        KH_HIP(hipMemcpy2DAsync(dev, sizeof(double) * 2 * g.nzp, host, sizeof(double) * 2 * g.nzh,
                                sizeof(double) * 2 * g.nzh, (size_t)rows, hipMemcpyHostToDevice, r->stream));
    } else {
        std::vector<float> tmp((size_t)g.nspec * 2, 0.f);
        for (long long row = 0; row < rows; row++)
            for (int k = 0; k < 2 * g.nzh; k++)
                tmp[(size_t)row * 2 * g.nzp + k] = (float)host[(size_t)row * 2 * g.nzh + k];
        KH_HIP(hipMemcpyAsync(dev, tmp.data(), sizeof(float) * tmp.size(), hipMemcpyHostToDevice, r->stream));
    }
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}
int download_spec(kh_run *r, double *host, const void *dev) {
    const Grid     &g    = r->g;
    const long long rows = (long long)g.nx * g.ny;
    if (!r->mixed) {
        KH_HIP(hipMemcpy2DAsync(host, sizeof(double) * 2 * g.nzh, dev, sizeof(double) * 2 * g.nzp,
                                sizeof(double) * 2 * g.nzh, (size_t)rows, hipMemcpyDeviceToHost, r->stream));
        KH_HIP(hipStreamSynchronize(r->stream));
    } else {
        std::vector<float> tmp((size_t)g.nspec * 2);
        KH_HIP(hipMemcpyAsync(tmp.data(), dev, sizeof(float) * tmp.size(), hipMemcpyDeviceToHost, r->stream));
        KH_HIP(hipStreamSynchronize(r->stream));
        for (long long row = 0; row < rows; row++)
            for (int k = 0; k < 2 * g.nzh; k++)
                host[(size_t)row * 2 * g.nzh + k] = tmp[(size_t)row * 2 * g.nzp + k];
    }
    return KH_OK;
}

void *field_buffer(kh_run *r, int field) {
    switch (field) {
        case KH_FIELD_U0: case KH_FIELD_U1: case KH_FIELD_U2: return r->ureal[field - KH_FIELD_U0];
        case KH_FIELD_UP0: case KH_FIELD_UP1: case KH_FIELD_UP2: return r->up[field - KH_FIELD_UP0];
        case KH_FIELD_PHI_SUM: return r->phisum;
        case KH_FIELD_F0: case KH_FIELD_F1: case KH_FIELD_F2: return r->work[field - KH_FIELD_F0];
        default: return nullptr;
    }
}

void drop_graph(kh_run *r);

int free_run(kh_run *r) {
    if (!r) return KH_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    r->stream = r->own_stream;  // a batched run goes back to its own stream
    drop_graph(r);
    void *bufs[] = {r->zeta[0], r->zeta[1], r->spec[0], r->spec[1], r->spec[2], r->spec[3], r->spec[4],
                    r->ureal[0], r->ureal[1], r->ureal[2], r->up[0], r->up[1], r->up[2], r->work[0], r->work[1],
                    r->work[2], r->work[3], r->work[4], r->phisum, r->ukdc, r->dp, r->ds, r->sek, r->cphi, r->cidx,
                    r->ccode, r->ccnt, r->d_ts, r->slipeps, r->slipdc, r->slipflag, r->xs};
    if (!r->arena)  // buffers carved from a batch arena stay with the arena
        for (void *b : bufs)
            if (b) (void)hipFreeAsync(b, r->stream);
    if (r->traj) (void)hipFreeAsync(r->traj, r->stream);
    if (r->frames) (void)hipFreeAsync(r->frames, r->stream);
    for (double *b : {r->shear_pstress, r->shear_acc, r->shear_last, r->shear_hist})
        if (b) (void)hipFreeAsync(b, r->stream);
    for (void *b : {(void *)r->ra.B, (void *)r->ra.GR, (void *)r->ra.GRb, (void *)r->bid, (void *)r->flo})
        if (b) (void)hipFreeAsync(b, r->stream);
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto &p : r->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : r->event_pool) (void)hipEventDestroy(e);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
    return KH_OK;
}

int check_params(const kh_params *P, int np) {
    if (!P) return fail(KH_ERR_ARG, "params is NULL");
    const int n[3] = {P->nx, P->ny, P->nz};
    for (int d = 0; d < 3; d++)
        if (!fft_length_supported(n[d]))
            return fail(KH_ERR_UNSUPPORTED, "grid size %d is not a supported FFT length (2^a 3^b, 8..512)", n[d]);
    if (!(P->dx > 0) || !(P->rho > 0) || !(P->eta > 0) || !(P->dt > 0) || !(P->a > 0) || !(P->a_xi > 0))
        return fail(KH_ERR_ARG, "dx, rho, eta, dt, a, a_xi must be positive");
    if (P->g_dir < 0 || P->g_dir > 2) return fail(KH_ERR_ARG, "g_dir must be 0, 1 or 2");
    if (P->nspecies < 1 || !P->species) return fail(KH_ERR_ARG, "at least one species is required");
    if (P->precision != KH_FP64 && P->precision != KH_MIXED) return fail(KH_ERR_ARG, "precision must be KH_FP64 or KH_MIXED");
    if (np < 0) return fail(KH_ERR_ARG, "negative particle count");
    if (P->lj_powers < 0 || P->lj_powers > 2) return fail(KH_ERR_UNSUPPORTED, "LJ_powers must be 12:6, 24:12 or 36:18");
    int nh = (int)P->a + (int)P->a_xi + 2;  // Particle_domain half-width (profile.cxx:34)
    if (nh > 100) return fail(KH_ERR_UNSUPPORTED, "particle radius too large for the Sekibun cell encoding");
    if (np > 0 && (nh >= P->nx || nh >= P->ny || nh >= P->nz))
        return fail(KH_ERR_ARG, "particle domain (half-width %d) does not fit the grid", nh);
    for (int i = 0; i < P->nspecies; i++) {
        if (P->species[i].propulsion == KH_JP_SLIP && P->slip_iter_max < 1)
            return fail(KH_ERR_ARG, "slip_iter_max must be at least 1 with slip species");
        if (!(P->species[i].mass_ratio > 0)) return fail(KH_ERR_ARG, "mass_ratio must be positive");
    }
    return KH_OK;
}

// Device bytes of one run (the allocations of kh_run_create).
size_t run_footprint(const kh_params &P, int np) {
    const size_t esz = P.precision == KH_MIXED ? 4 : 8;
    const size_t nzh = (size_t)P.nz / 2 + 1, per = 128 / (2 * esz), nzp = (nzh + per - 1) / per * per;
    const size_t nreal = (size_t)P.nx * P.ny * P.nz, nspec = (size_t)P.nx * P.ny * nzp;
    const int    nh = (int)P.a + (int)P.a_xi + 2;
    const double rmax = P.a * P.dx + 0.5 * P.a_xi * P.dx + 1.7321 * P.dx;
    size_t       npts = 0;
    for (int i = -nh; i <= nh; i++)
        for (int j = -nh; j <= nh; j++)
            for (int k = -nh; k <= nh; k++)
                if (sqrt((i * P.dx) * (i * P.dx) + (j * P.dx) * (j * P.dx) + (k * P.dx) * (k * P.dx)) < rmax) npts++;
    const size_t npp = (size_t)(np > 0 ? np : 1);
    return 7 * 2 * esz * nspec + 12 * esz * nreal + npp * (sizeof(DevParticle) + 4 * sizeof(double)) +
           npp * npts * (esz + 2 * sizeof(int)) + npp * sizeof(int) +
           npts * sizeof(int) + (size_t)P.nspecies * sizeof(DevSpecies) + 1024;
}

void drop_graph(kh_run *r) {
    if (r->graph_exec) (void)hipGraphExecDestroy(r->graph_exec);
    if (r->graph) (void)hipGraphDestroy(r->graph);
    if (r->graph_exec_k) (void)hipGraphExecDestroy(r->graph_exec_k);
    if (r->graph_k) (void)hipGraphDestroy(r->graph_k);
    r->graph_exec   = nullptr;
    r->graph        = nullptr;
    r->graph_exec_k = nullptr;
    r->graph_k      = nullptr;
}

}  // namespace

// ------------------------------------------------------ rigid bodies (host)
// The initial block of a rigid run (init_particle.cxx:455-487) in FP64 on
// the host, once: its sums run in the reference's order (beads in order,
// each over its Sekibun points), so the principal frame it finds is the
// reference's.
namespace rigid_init {

struct Geo {
    int    n[3];
    double dx, L[3], radius, hxi;
};
int    nint_h(double a) { return a >= 0.0 ? (int)(a + .5) : (int)(a - .5); }
// Distance (interaction.h:125-132): minimum-image |x2 - x1|
double dist_h(const Geo &g, const double *x1, const double *x2) {
    double s = 0.0;
    for (int d = 0; d < 3; d++) {
        double e = x2[d] - x1[d];
        e -= (double)nint_h(e / g.L[d]) * g.L[d];
        s += e * e;
    }
    return sqrt(s);
}
// Phi (profile.h:30-50)
double phi_h(const Geo &g, double x) {
    auto H = [&](double y) { return y > 0 ? exp(-(g.dx / y) * (g.dx / y)) : 0.0; };
    const double a1 = g.radius + g.hxi - x, a2 = x - g.radius + g.hxi;
    const double h1 = H(a1);
    return h1 / (h1 + H(a2));
}
// Particle_cell + Relative_coord (make_phi.h:114-170) for one Sekibun offset:
// grid index, r = cell dx - residue, grid position
struct Pt {
    long long idx;
    double    r[3], x[3];
};
Pt point_h(const Geo &g, const double *xp, int code) {
    const int cell[3] = {(code & 255) - 128, ((code >> 8) & 255) - 128, ((code >> 16) & 255) - 128};
    Pt        p;
    int       rm[3];
    for (int d = 0; d < 3; d++) {
        const double dmy   = xp[d] / g.dx;
        const int    x_int = (int)dmy;
        const double res   = (dmy - (double)x_int) * g.dx;
        rm[d]              = (x_int + cell[d] + g.n[d]) % g.n[d];
        p.r[d]             = (double)cell[d] * g.dx - res;
        p.x[d]             = rm[d] * g.dx;
    }
    p.idx = ((long long)rm[0] * g.n[1] + rm[1]) * g.n[2] + rm[2];
    return p;
}
void cross(double *c, const double *a, const double *b) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
// qtn_qvq_prod (quaternion.h:178-185) with q (BODY2SPACE) or its conjugate
void rotate(double *o, const double *q, const double *v, bool to_body) {
    const double s = q[0], a = to_body ? -q[1] : q[1], b = to_body ? -q[2] : q[2], c = to_body ? -q[3] : q[3];
    const double qs2 = s * s - (a * a + b * b + c * c), qdotp = a * v[0] + b * v[1] + c * v[2];
    o[0] = (qs2 * v[0] + 2.0 * (qdotp * a + s * (b * v[2] - c * v[1])));
    o[1] = (qs2 * v[1] + 2.0 * (qdotp * b + s * (c * v[0] - a * v[2])));
    o[2] = (qs2 * v[2] + 2.0 * (qdotp * c + s * (a * v[1] - b * v[0])));
}
// Cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (the scheme
// of Numerical Recipes section 11.1 as jacobi, matrix_diagonal.cxx:56-172,
// applies it: entries scaled by their largest magnitude, thresholds for the
// first three sweeps, negligible elements dropped after four); eigenvectors
// returned as rows, by decreasing eigenvalue (jacobi_ordered, 25-53).
bool jacobi(const double In[3][3], double V[3][3], double ev[3]) {
    double A[3][3], b[3], z[3], mval = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) mval = std::max(mval, fabs(In[i][j]));
    const double minv = mval > 0.0 ? 1.0 / mval : 1.0;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            A[i][j] = In[i][j] * minv;
            V[i][j] = i == j ? 1.0 : 0.0;
        }
        b[i] = ev[i] = A[i][i];
        z[i]         = 0.0;
    }
    auto rot = [](double M[3][3], double sn, double tau, int i, int j, int k, int l) {
        const double g = M[i][j], h = M[k][l];
        M[i][j]        = g - sn * (h + g * tau);
        M[k][l]        = h + sn * (g - h * tau);
    };
    const double tol = 10.0 * 2.220446049250313e-16;  // TOL_MP (macro.h:36-37)
    for (int it = 1; it <= 50; it++) {
        const double sm = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
        if (!(sm > tol || -tol > sm)) {
            for (int i = 0; i < 2; i++)
                for (int j = i + 1; j < 3; j++) std::swap(V[i][j], V[j][i]);
            for (int i = 0; i < 2; i++) {  // by decreasing eigenvalue, the last of equals
                int    k   = i;
                double top = ev[k];
                for (int j = i; j < 3; j++)
                    if (ev[j] >= top) {
                        k   = j;
                        top = ev[k];
                    }
                if (k != i) {
                    ev[k] = ev[i];
                    ev[i] = top;
                    for (int j = 0; j < 3; j++) std::swap(V[k][j], V[i][j]);
                }
            }
            for (int i = 0; i < 3; i++) ev[i] *= mval;
            return true;
        }
        const double tresh = it < 4 ? 0.2 * sm / 9.0 : 0.0;
        for (int ip = 0; ip < 2; ip++)
            for (int iq = ip + 1; iq < 3; iq++) {
                const double g = 100.0 * fabs(A[ip][iq]);
                if (it > 4 && g <= 1.0e-17 * fabs(ev[ip]) && g <= 1.0e-17 * fabs(ev[iq])) {
                    A[ip][iq] = 0.0;
                } else if (fabs(A[ip][iq]) > tresh) {
                    double h = ev[iq] - ev[ip], t;
                    if (g <= 1.0e-17 * fabs(h)) {
                        t = A[ip][iq] / h;
                    } else {
                        const double theta = 0.5 * h / A[ip][iq];
                        t                  = 1.0 / (fabs(theta) + sqrt(1.0 + theta * theta));
                        if (theta < 0.0) t = -t;
                    }
                    const double c = 1.0 / sqrt(1 + t * t), sn = t * c, tau = sn / (1.0 + c);
                    h = t * A[ip][iq];
                    z[ip] -= h;
                    z[iq] += h;
                    ev[ip] -= h;
                    ev[iq] += h;
                    A[ip][iq] = 0.0;
                    for (int j = 0; j < ip; j++) rot(A, sn, tau, j, ip, j, iq);
                    for (int j = ip + 1; j < iq; j++) rot(A, sn, tau, ip, j, j, iq);
                    for (int j = iq + 1; j < 3; j++) rot(A, sn, tau, ip, j, iq, j);
                    for (int j = 0; j < 3; j++) rot(V, sn, tau, j, ip, j, iq);
                }
            }
        for (int i = 0; i < 3; i++) {
            b[i] += z[i];
            ev[i] = b[i];
            z[i]  = 0.0;
        }
    }
    return false;
}
// The body frame from its principal axes: ex along the axis of the
// smallest moment, ey the largest's made orthogonal to it, ez = ex x ey
// (M_coordinate_frame, lad3.h:318-354); rows V[2] <- ex, V[0] <- ey, V[1] <- ez
void frame(double V[3][3]) {
    double ex[3], ey[3], ez[3], *u = V[2], *v = V[0];
    double nrm = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    for (int i = 0; i < 3; i++) ex[i] = u[i] / nrm;
    const double dot = ex[0] * v[0] + ex[1] * v[1] + ex[2] * v[2];
    for (int i = 0; i < 3; i++) ey[i] = v[i] - dot * ex[i];
    nrm = sqrt(ey[0] * ey[0] + ey[1] * ey[1] + ey[2] * ey[2]);
    for (int i = 0; i < 3; i++) ey[i] /= nrm;
    cross(ez, ex, ey);
    nrm = sqrt(ez[0] * ez[0] + ez[1] * ez[1] + ez[2] * ez[2]);
    for (int i = 0; i < 3; i++) {
        V[2][i] = ex[i];
        V[0][i] = ey[i];
        V[1][i] = ez[i] / nrm;
    }
}
// Rotation matrix -> unit quaternion (s, v) from its largest q_i^2
// component (Shepperd's method; rm_rqtn, quaternion.h:491-549)
void to_quaternion(const double R[3][3], double *q) {
    double qq[4] = {1.0 + R[0][0] + R[1][1] + R[2][2], 1.0 + R[0][0] - R[1][1] - R[2][2],
                    1.0 - R[0][0] + R[1][1] - R[2][2], 1.0 - R[0][0] - R[1][1] + R[2][2]};
    int    i     = 0;
    for (int d = 1; d < 4; d++)
        if (qq[d] > qq[i]) i = d;
    qq[i]            = sqrt(qq[i]) / 2.0;
    const double dmy = 1.0 / (4.0 * qq[i]);
    const double sk[3] = {R[2][1] - R[1][2], R[0][2] - R[2][0], R[1][0] - R[0][1]};  // 4 s v
    const double sy[3] = {R[1][0] + R[0][1], R[0][2] + R[2][0], R[2][1] + R[1][2]};  // 4 v_a v_b
    if (i == 0) {
        qq[1] = sk[0] * dmy;
        qq[2] = sk[1] * dmy;
        qq[3] = sk[2] * dmy;
    } else if (i == 1) {
        qq[0] = sk[0] * dmy;
        qq[2] = sy[0] * dmy;
        qq[3] = sy[1] * dmy;
    } else if (i == 2) {
        qq[0] = sk[1] * dmy;
        qq[1] = sy[0] * dmy;
        qq[3] = sy[2] * dmy;
    } else {
        qq[0] = sk[2] * dmy;
        qq[1] = sy[1] * dmy;
        qq[2] = sy[2] * dmy;
    }
    for (int k = 0; k < 4; k++) q[k] = qq[k];
}

}  // namespace rigid_init

// ================================================================== C ABI
extern "C" {

const char *kh_last_error(void) { return g_last_error.c_str(); }

const char *kh_strerror(int s) {
    switch (s) {
        case KH_OK: return "ok";
        case KH_ERR_ARG: return "invalid argument";
        case KH_ERR_HIP: return "HIP runtime error";
        case KH_ERR_NOMEM: return "device allocation failed";
        case KH_ERR_UNSUPPORTED: return "unsupported";
        case KH_ERR_STATE: return "invalid state";
        case KH_ERR_NONFINITE: return "non-finite particle state";
        default: return "unknown status";
    }
}

}  // extern "C"

// kh_run_create; with an arena, the run's device buffers are carved from it
// (in the same order and sizes for every run of one grid and precision, so
// the field buffers sit at the same offsets in every arena) and not freed
// with the run.
int run_create(const kh_params *P, int np, char *arena, size_t arena_bytes, kh_run **out) {
    if (!out) return fail(KH_ERR_ARG, "out is NULL");
    *out = nullptr;
    KH_TRY(check_params(P, np));
    kh_run *r = new kh_run();
    r->arena  = arena;
    r->P      = *P;
    r->species.assign(P->species, P->species + P->nspecies);
    r->P.species = r->species.data();
    r->np        = np;
    r->mixed     = P->precision == KH_MIXED;
    r->esz       = r->mixed ? 4 : 8;
    int st       = KH_OK;
    auto bail    = [&](int s) {
        free_run(r);
        return s;
    };
    if (hipGetDevice(&r->device) != hipSuccess) return bail(fail(KH_ERR_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(KH_ERR_HIP, "hipStreamCreate failed"));
    r->own_stream = r->stream;

    // Set_global_parameters (input.cxx:424-600)
    Grid &g = r->g;
    g.nx = P->nx;
    g.ny = P->ny;
    g.nz = P->nz;
    g.nzh = P->nz / 2 + 1;
    {  // spectral rows start on 128-byte boundaries
        const int per = 128 / (2 * (P->precision == KH_MIXED ? 4 : 8));
        g.nzp         = (g.nzh + per - 1) / per * per;
    }
    g.hnx = P->nx / 2;
    g.hny = P->ny / 2;
    g.trn_x = (P->nx + 2) / 3;
    g.trn_y = (P->ny + 2) / 3;
    g.trn_z = (P->nz + 2) / 3;
    g.wave_x = 2. * KH_PI / (P->nx * P->dx);
    g.wave_y = 2. * KH_PI / (P->ny * P->dx);
    g.wave_z = 2. * KH_PI / (P->nz * P->dx);
    g.nreal  = (long long)g.nx * g.ny * g.nz;
    g.nspec  = (long long)g.nx * g.ny * g.nzp;
    r->radius = P->a * P->dx;
    r->xi     = P->a_xi * P->dx;
    r->hxi    = 0.5 * r->xi;
    r->rho    = P->rho;
    r->eta    = P->eta;
    r->nu     = P->eta / P->rho;
    r->dx3    = P->dx * P->dx * P->dx;
    r->kbt    = P->kbt;
    const double sigma = 2.0 * r->radius;
    double       a_r_cutoff;
    if (P->lj_truncate > 0)
        a_r_cutoff = (P->lj_powers == 0) ? pow(2., 1. / 6.) : (P->lj_powers == 1) ? pow(2., 1. / 12.) : pow(2., 1. / 18.);
    else if (P->lj_truncate == 0) {
        int nmin   = P->nx < P->ny ? P->nx : P->ny;
        nmin       = nmin < P->nz ? nmin : P->nz;
        a_r_cutoff = fmin(nmin * P->dx * .5 / sigma, 2.5);
    } else
        a_r_cutoff = 0.;
    r->lj_cutoff        = a_r_cutoff * sigma;
    r->gravity_on_fluid = P->g * P->rho * 4. / 3. * KH_PI * (r->radius * r->radius) * r->radius;
    r->fixmask          = (P->fix_cell[0] ? 1 : 0) | (P->fix_cell[1] ? 2 : 0) | (P->fix_cell[2] ? 4 : 0);
    std::vector<DevSpecies> hs(P->nspecies);
    for (int i = 0; i < P->nspecies; i++) {
        const kh_species &s     = P->species[i];
        const double      m     = s.mass_ratio * P->rho * 4. / 3. * KH_PI * (r->radius * r->radius * r->radius);
        DevSpecies       &d     = hs[i];
        d.mass_ratio            = s.mass_ratio;
        d.imass                 = 1. / m;
        d.imoi                  = 1. / (2. / 5. * m * (r->radius * r->radius));
        for (int k = 0; k < 3; k++) {
            d.janus_force[k]  = s.janus_force[k];
            d.janus_torque[k] = s.janus_torque[k];
        }
        d.slip_vel       = s.slip_vel;
        d.slip_mode_half = s.slip_mode / 2.0;
        d.c1             = s.rotlet_c1;
        d.c2             = s.rotlet_c2;
        d.propulsion     = s.propulsion;
        d.janus_axis     = s.janus_axis;
        if (s.propulsion == KH_JP_MOTOR) r->motor = 1;
        if (s.propulsion == KH_JP_SLIP) r->slip = 1;
    }
    // Particle_domain (profile.cxx:22-65)
    std::vector<int> sek;
    {
        const int    nh         = (int)P->a + (int)P->a_xi + 2;
        const double max_radius = r->radius + r->hxi + 1.7321 * P->dx;
        for (int i = -nh; i <= nh; i++)
            for (int j = -nh; j <= nh; j++)
                for (int k = -nh; k <= nh; k++) {
                    double dmy = sqrt((i * P->dx) * (i * P->dx) + (j * P->dx) * (j * P->dx) + (k * P->dx) * (k * P->dx));
                    if (dmy < max_radius) sek.push_back(sek_pack(i, j, k));
                }
        r->np_domain = (int)sek.size();
        r->sek_host  = sek;
    }

    // device memory
    const size_t rb = r->esz * (size_t)g.nreal, sb = 2 * r->esz * (size_t)g.nspec;
    // stream-ordered allocations: creating or destroying a run does not
    // synchronise the device, so runs on other streams keep going
    auto alloc = [&](void **p, size_t bytes) -> int {
        if (arena) {  // 256-byte aligned, in allocation order
            const size_t off = (r->arena_used + 255) / 256 * 256;
            if (off + bytes > arena_bytes) {
                *p = nullptr;
                return fail(KH_ERR_NOMEM, "run needs more than its %zu-byte arena", arena_bytes);
            }
            *p            = arena + off;
            r->arena_used = off + bytes;
            return KH_OK;
        }
        if (hipMallocAsync(p, bytes, r->stream) != hipSuccess) {
            *p = nullptr;
            return fail(KH_ERR_NOMEM, "device allocation of %zu bytes failed", bytes);
        }
        return KH_OK;
    };
    for (int i = 0; i < 2 && st == KH_OK; i++) st = alloc(&r->zeta[i], sb);
    for (int i = 0; i < 5 && st == KH_OK; i++) st = alloc(&r->spec[i], sb);
    for (int i = 0; i < 3 && st == KH_OK; i++) st = alloc(&r->ureal[i], rb);
    for (int i = 0; i < 3 && st == KH_OK; i++) st = alloc(&r->up[i], rb);
    for (int i = 0; i < 5 && st == KH_OK; i++) st = alloc(&r->work[i], rb);
    if (st == KH_OK) st = alloc(&r->phisum, rb);
    r->occw = ((g.nz + 3) / 4 + 63) / 64;
    if (st == KH_OK) st = alloc(&r->occ, sizeof(unsigned long long) * (size_t)r->occw * (size_t)g.nx * g.ny);
    if (st == KH_OK) st = alloc((void **)&r->ukdc, 3 * sizeof(double));
    if (st == KH_OK) st = alloc((void **)&r->d_ts, sizeof(int));
    if (st == KH_OK) st = alloc((void **)&r->slipeps, sizeof(double));
    if (st == KH_OK) st = alloc((void **)&r->xs, 4 * sizeof(double) * (size_t)(np > 0 ? np : 1));
    if (st == KH_OK) st = alloc((void **)&r->slipdc, 3 * sizeof(double));
    if (st == KH_OK) st = alloc((void **)&r->slipflag, 32);
    if (st == KH_OK) st = alloc((void **)&r->dp, sizeof(DevParticle) * (size_t)(np > 0 ? np : 1));
    if (st == KH_OK) st = alloc((void **)&r->ds, sizeof(DevSpecies) * hs.size());
    if (st == KH_OK) st = alloc((void **)&r->sek, sizeof(int) * sek.size());
    const size_t ncache = (size_t)(np > 0 ? np : 1) * sek.size();
    if (st == KH_OK) st = alloc(&r->cphi, r->esz * ncache);
    if (st == KH_OK) st = alloc((void **)&r->cidx, sizeof(int) * ncache);
    if (st == KH_OK) st = alloc((void **)&r->ccode, sizeof(int) * ncache);
    if (st == KH_OK) st = alloc((void **)&r->ccnt, sizeof(int) * (size_t)(np > 0 ? np : 1));
    if (st != KH_OK) return bail(st);
    hipStream_t  cs  = r->stream;
    const size_t psz = sizeof(DevParticle) * (size_t)(np > 0 ? np : 1);
    if (hipMemcpyAsync(r->ds, hs.data(), sizeof(DevSpecies) * hs.size(), hipMemcpyHostToDevice, cs) != hipSuccess ||
        hipMemcpyAsync(r->sek, sek.data(), sizeof(int) * sek.size(), hipMemcpyHostToDevice, cs) != hipSuccess ||
        hipMemsetAsync(r->ukdc, 0, 3 * sizeof(double), cs) != hipSuccess ||
        hipMemsetAsync(r->d_ts, 0, sizeof(int), cs) != hipSuccess || hipMemsetAsync(r->dp, 0, psz, cs) != hipSuccess ||
        hipMemsetAsync(r->slipflag, 0, 32, cs) != hipSuccess)
        return bail(fail(KH_ERR_HIP, "initial upload failed"));
    for (int i = 0; i < 2; i++)
        if (hipMemsetAsync(r->zeta[i], 0, sb, cs) != hipSuccess) return bail(fail(KH_ERR_HIP, "hipMemset failed"));
    for (int i = 0; i < 5; i++)  // the row padding stays zero for good
        if (hipMemsetAsync(r->spec[i], 0, sb, cs) != hipSuccess) return bail(fail(KH_ERR_HIP, "hipMemset failed"));
    void *zr[] = {r->ureal[0], r->ureal[1], r->ureal[2], r->up[0], r->up[1], r->up[2],
                  r->work[0], r->work[1], r->work[2], r->work[3], r->work[4], r->phisum};
    for (void *b : zr)
        if (hipMemsetAsync(b, 0, rb, cs) != hipSuccess) return bail(fail(KH_ERR_HIP, "hipMemset failed"));
    if (hipMemsetAsync(r->occ, 0, sizeof(unsigned long long) * (size_t)r->occw * (size_t)g.nx * g.ny, cs) != hipSuccess)
        return bail(fail(KH_ERR_HIP, "hipMemset failed"));
    if (hipStreamSynchronize(cs) != hipSuccess) return bail(fail(KH_ERR_HIP, "initial upload failed"));

    PParams &pp        = r->pp;
    pp.nx              = g.nx;
    pp.ny              = g.ny;
    pp.nz              = g.nz;
    pp.dx              = P->dx;
    pp.L[0]            = g.nx * P->dx;
    pp.L[1]            = g.ny * P->dx;
    pp.L[2]            = g.nz * P->dx;
    pp.radius          = r->radius;
    pp.hxi             = r->hxi;
    pp.xi              = r->xi;
    pp.rotation        = P->rotation;
    pp.np              = np;
    pp.np_domain       = r->np_domain;
    pp.sek             = r->sek;
    pp.species         = r->ds;
    pp.lj_on           = P->lj_truncate >= 0;
    pp.lj_powers       = P->lj_powers;
    pp.lj_cell_compat  = P->lj_cell_compat;
    pp.lj_dia          = sigma;
    pp.lj_cutoff       = r->lj_cutoff;
    pp.epsilon         = P->epsilon;
    pp.g               = P->g;
    pp.g_dir           = P->g_dir;
    pp.gravity_on_fluid = r->gravity_on_fluid;
    pp.motor           = r->motor;
    pp.xs              = r->xs;
    pp.rho             = P->rho;
    {
        const int nh    = (int)P->a + (int)P->a_xi + 2, nmin = std::min(P->nx, std::min(P->ny, P->nz));
        pp.direct_dist = 2 * (nh + 2) <= nmin;
    }
    // squirmer steps decide the slip loop on the device where the wavefront
    // z engine and its masked inverse pass take NZ
    r->slip_dev        = r->slip && zw_supported(g.nz) && g.nz >= 32;
    r->zclear          = (!r->slip || r->slip_dev) && z_force_clears(g);

    CTime t0  = {0, 0.0, P->dt, P->dt * 0.5, P->dt, P->dt * 0.5};  // sp_3d_ns.cxx:782
    r->jikan  = t0;
    *out      = r;
    return KH_OK;
}

extern "C" {

int kh_run_create(const kh_params *P, int np, kh_run **out) { return run_create(P, np, nullptr, 0, out); }

int kh_run_destroy(kh_run *r) { return free_run(r); }

int kh_set_particles(kh_run *r, const Particle *p, int n) {
    if (!r || (!p && n > 0)) return fail(KH_ERR_ARG, "NULL argument");
    if (n != r->np) return fail(KH_ERR_ARG, "particle count %d != run's %d", n, r->np);
    std::vector<DevParticle> h((size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        const Particle &s = p[i];
        DevParticle    &d = h[i];
        memset(&d, 0, sizeof d);
        if (s.spec < 0 || s.spec >= r->P.nspecies) return fail(KH_ERR_ARG, "particle %d: bad species %d", i, s.spec);
        const double L[3] = {r->pp.L[0], r->pp.L[1], r->pp.L[2]};
        for (int k = 0; k < 3; k++) {
            if (!(s.x[k] >= 0 && s.x[k] < L[k])) return fail(KH_ERR_ARG, "particle %d outside the box", i);
            d.x[k] = s.x[k];
            d.x_prev[k] = s.x_previous[k];
            d.x_nopbc[k] = s.x_nopbc[k];
            d.v[k] = s.v[k];
            d.v_old[k] = s.v_old[k];
            d.om[k] = s.omega[k];
            d.om_old[k] = s.omega_old[k];
            d.fr[k] = s.fr[k];
            d.fr_prev[k] = s.fr_previous[k];
            d.tr[k] = s.torque_r[k];
            d.tr_prev[k] = s.torque_r_previous[k];
            d.fh[k] = s.f_hydro[k];
            d.fh_prev[k] = s.f_hydro_previous[k];
            d.th[k] = s.torque_hydro[k];
            d.th_prev[k] = s.torque_hydro_previous[k];
            d.fs[k] = s.f_slip[k];
            d.fs_prev[k] = s.f_slip_previous[k];
            d.ts[k] = s.torque_slip[k];
            d.ts_prev[k] = s.torque_slip_previous[k];
            d.v_slip[k] = s.v_slip[k];
            d.om_slip[k] = s.omega_slip[k];
            d.mass_center[k] = s.mass_center[k];
            d.smc[k] = s.surface_mass_center[k];
            d.sdv[k] = s.surface_dv[k];
            d.sdw[k] = s.surface_dw[k];
            d.mdfr[k] = s.momentum_depend_fr[k];
            for (int l = 0; l < 3; l++) {
                d.inertia[3 * k + l]  = s.inertia[k][l];
                d.sinertia[3 * k + l] = s.surface_inertia[k][l];
            }
        }
        d.mass  = s.mass;
        d.smass = s.surface_mass;
        d.q[0] = s.q.s;
        d.q_old[0] = s.q_old.s;
        for (int k = 0; k < 3; k++) {
            d.q[1 + k] = s.q.v[k];
            d.q_old[1 + k] = s.q_old.v[k];
        }
        d.spec = s.spec;
    }
    r->pcache_valid = 0;
    if (r->rigid) {  // new particles: their bodies are set again with kh_set_rigid
        r->rigid  = 0;
        r->pp.bid = nullptr;
        r->pp.flo = nullptr;
        drop_graph(r);
    }
    std::vector<double> xs(4 * (size_t)(n > 0 ? n : 1), 0.0);
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) xs[4 * i + k] = h[i].x[k];
        xs[4 * i + 3] = r->species[h[i].spec].propulsion == KH_JP_OBSTACLE ? 1.0 : 0.0;
    }
    if (n > 0) KH_HIP(hipMemcpyAsync(r->dp, h.data(), sizeof(DevParticle) * n, hipMemcpyHostToDevice, r->stream));
    if (n > 0) KH_HIP(hipMemcpyAsync(r->xs, xs.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}

int kh_get_particles(kh_run *r, Particle *p, int n) {
    if (!r || (!p && n > 0)) return fail(KH_ERR_ARG, "NULL argument");
    if (n != r->np) return fail(KH_ERR_ARG, "particle count %d != run's %d", n, r->np);
    std::vector<DevParticle> h((size_t)(n > 0 ? n : 1));
    if (n > 0) KH_HIP(hipMemcpyAsync(h.data(), r->dp, sizeof(DevParticle) * n, hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    for (int i = 0; i < n; i++) {
        Particle          &s = p[i];
        const DevParticle &d = h[i];
        for (int k = 0; k < 3; k++) {
            s.x[k] = d.x[k];
            s.x_previous[k] = d.x_prev[k];
            s.x_nopbc[k] = d.x_nopbc[k];
            s.v[k] = d.v[k];
            s.v_old[k] = d.v_old[k];
            s.omega[k] = d.om[k];
            s.omega_old[k] = d.om_old[k];
            s.fr[k] = d.fr[k];
            s.fr_previous[k] = d.fr_prev[k];
            s.torque_r[k] = d.tr[k];
            s.torque_r_previous[k] = d.tr_prev[k];
            s.f_hydro[k] = d.fh[k];
            s.f_hydro_previous[k] = d.fh_prev[k];
            s.torque_hydro[k] = d.th[k];
            s.torque_hydro_previous[k] = d.th_prev[k];
            s.f_slip[k] = d.fs[k];
            s.f_slip_previous[k] = d.fs_prev[k];
            s.torque_slip[k] = d.ts[k];
            s.torque_slip_previous[k] = d.ts_prev[k];
            s.v_slip[k] = d.v_slip[k];
            s.omega_slip[k] = d.om_slip[k];
            s.mass_center[k] = d.mass_center[k];
            s.surface_mass_center[k] = d.smc[k];
            s.surface_dv[k] = d.sdv[k];
            s.surface_dw[k] = d.sdw[k];
            s.momentum_depend_fr[k] = d.mdfr[k];
            for (int l = 0; l < 3; l++) {
                s.inertia[k][l]         = d.inertia[3 * k + l];
                s.surface_inertia[k][l] = d.sinertia[3 * k + l];
            }
        }
        s.mass         = d.mass;
        s.surface_mass = d.smass;
        s.q.s = d.q[0];
        s.q_old.s = d.q_old[0];
        for (int k = 0; k < 3; k++) {
            s.q.v[k] = d.q[1 + k];
            s.q_old.v[k] = d.q_old[1 + k];
        }
        s.spec = d.spec;
    }
    return KH_OK;
}

// Forms zeta's modes that a sheared step left to be formed (Step::settle):
// every entry point that reads zeta, changes ureal or the strain, or runs
// anything but the sheared step calls it first.
static int settle_zeta(kh_run *r) {
    if (!r->zeta_lazy) return KH_OK;
    return r->mixed ? Step<float>::settle(r) : Step<double>::settle(r);
}

int kh_set_fluid(kh_run *r, const double *z0, const double *z1, const double uk_dc[3]) {
    if (!r || !z0 || !z1 || !uk_dc) return fail(KH_ERR_ARG, "NULL argument");
    r->zeta_lazy = 0;  // every mode is overwritten
    KH_TRY(upload_spec(r, r->zeta[0], z0));
    KH_TRY(upload_spec(r, r->zeta[1], z1));
    r->zeta_dirty = 1;
    KH_HIP(hipMemcpy(r->ukdc, uk_dc, 3 * sizeof(double), hipMemcpyHostToDevice));
    return KH_OK;
}

int kh_get_fluid(kh_run *r, double *z0, double *z1, double uk_dc[3]) {
    if (!r || !z0 || !z1 || !uk_dc) return fail(KH_ERR_ARG, "NULL argument");
    KH_TRY(settle_zeta(r));
    KH_TRY(download_spec(r, z0, r->zeta[0]));
    KH_TRY(download_spec(r, z1, r->zeta[1]));
    KH_HIP(hipMemcpy(uk_dc, r->ukdc, 3 * sizeof(double), hipMemcpyDeviceToHost));
    return KH_OK;
}

int kh_set_time(kh_run *r, const CTime *t) {
    if (!r || !t) return fail(KH_ERR_ARG, "NULL argument");
    if (t->dt_fluid != r->jikan.dt_fluid || t->dt_md != r->jikan.dt_md) {
        drop_graph(r);  // dt enters the captured kernel arguments
    }
    r->jikan = *t;
    KH_HIP(hipMemcpyAsync(r->d_ts, &r->jikan.ts, sizeof(int), hipMemcpyHostToDevice, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}
int kh_get_time(kh_run *r, CTime *t) {
    if (!r || !t) return fail(KH_ERR_ARG, "NULL argument");
    *t = r->jikan;
    return KH_OK;
}

int kh_get_field(kh_run *r, int field, double *out) {
    if (!r || !out) return fail(KH_ERR_ARG, "NULL argument");
    KH_TRY(settle_zeta(r));
    if (field == KH_FIELD_PHI) {  // phi = min(phi_sum, 1) (make_phi.cxx:248-259)
        KH_TRY(download_real(r, out, r->phisum));
        const size_t n = (size_t)r->g.nx * r->g.ny * 2 * r->g.nzh;
        for (size_t i = 0; i < n; i++) out[i] = fmin(out[i], 1.0);
        return KH_OK;
    }
    void *b = field_buffer(r, field);
    if (!b) return fail(KH_ERR_ARG, "unknown field %d", field);
    return download_real(r, out, b);
}

int kh_set_field(kh_run *r, int field, const double *in) {
    if (!r || !in) return fail(KH_ERR_ARG, "NULL argument");
    KH_TRY(settle_zeta(r));
    r->pclean = false;
    void *b = field_buffer(r, field);
    if (!b) return fail(KH_ERR_ARG, "field %d cannot be written", field);
    return upload_real(r, b, in);
}

int kh_init_fluid(kh_run *r) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    r->zeta_lazy = 0;  // every mode is overwritten
    r->zeta_dirty = 1;
    return r->mixed ? Step<float>::init_fluid(r) : Step<double>::init_fluid(r);
}

// Steady-state steps (ts >= 1, fused path, zeta clean) are identical up to
// the device step counter, so `count` of them captured back to back form
// one HIP graph; replays then cost one launch per `count` steps. Small grids
// (the serving loop) are bound by launch cost, hence the multi-step graph.
constexpr int GRAPH_STEPS = 8;
static int capture_steps(kh_run *r, int count, hipGraph_t *graph, hipGraphExec_t *exec) {
    KH_HIP(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
    r->capturing = true;
    int st       = KH_OK;
    for (int k = 0; k < count && st == KH_OK; k++) st = r->mixed ? Step<float>::step(r) : Step<double>::step(r);
    r->capturing = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(r->stream, &g);
    if (st != KH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(KH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    *graph = g;
    KH_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return KH_OK;
}

int kh_step(kh_run *r, int nsteps) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (nsteps < 0) return fail(KH_ERR_ARG, "negative step count");
    if (!r->shear) KH_TRY(settle_zeta(r));
    for (int s = 0; s < nsteps;) {
        // (a slip loop decided on the host is not captured)
        const bool graph_ok = r->use_graph && !r->timing && r->step_path == KH_PATH_FUSED && r->jikan.ts > 0 &&
                              !r->zeta_dirty && (!r->slip || r->slip_dev) && !r->shear &&
                              (!r->zclear || r->pclean);
        int done = 1;
        if (graph_ok && nsteps - s >= GRAPH_STEPS) {
            if (!r->graph_exec_k) KH_TRY(capture_steps(r, GRAPH_STEPS, &r->graph_k, &r->graph_exec_k));
            KH_HIP(hipGraphLaunch(r->graph_exec_k, r->stream));
            r->pcache_valid = 1;
            done = GRAPH_STEPS;
        } else if (graph_ok) {
            if (!r->graph_exec) KH_TRY(capture_steps(r, 1, &r->graph, &r->graph_exec));
            KH_HIP(hipGraphLaunch(r->graph_exec, r->stream));
            r->pcache_valid = 1;
        } else {
            KH_TRY(r->mixed ? Step<float>::step(r) : Step<double>::step(r));
        }
        for (int k = 0; k < done; k++) {
            r->jikan.time += r->jikan.dt_fluid;  // sp_3d_ns.cxx:955
            r->jikan.ts++;
        }
        s += done;
    }
    return KH_OK;
}

int kh_set_shear(kh_run *r, double shear_rate, int ac, double frequency) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (!std::isfinite(shear_rate) || !std::isfinite(frequency)) return fail(KH_ERR_ARG, "non-finite shear parameter");
    if (r->slip) return fail(KH_ERR_UNSUPPORTED, "the sheared step has no squirmer slip iteration");
    KH_TRY(settle_zeta(r));  // at the strain the last step left
    drop_graph(r);
    if (!r->shear_last) {  // the stress record's device buffers
        const size_t np = (size_t)(r->np > 0 ? r->np : 1);
        r->shear_nacc   = obl_to_cart_blocks(r->g);
        if (hipMallocAsync((void **)&r->shear_pstress, sizeof(double) * 6 * np, r->stream) != hipSuccess ||
            hipMallocAsync((void **)&r->shear_acc, sizeof(double) * r->shear_nacc, r->stream) != hipSuccess ||
            hipMallocAsync((void **)&r->shear_last, sizeof(double) * KH_SHEAR_RECORD, r->stream) != hipSuccess)
            return fail(KH_ERR_NOMEM, "stress record buffers");
        KH_HIP(hipMemsetAsync(r->shear_pstress, 0, sizeof(double) * 6 * np, r->stream));
        KH_HIP(hipMemsetAsync(r->shear_acc, 0, sizeof(double) * r->shear_nacc, r->stream));
        KH_HIP(hipMemsetAsync(r->shear_last, 0, sizeof(double) * KH_SHEAR_RECORD, r->stream));
        r->pp.pstress = r->shear_pstress;
    }
    r->strain     = 0.0;
    r->shear      = 1;
    r->shear_ac   = ac ? 1 : 0;
    r->shear_rate = shear_rate;
    r->shear_freq = frequency;
    r->xyaspect   = (double)r->g.nx / (double)r->g.ny;  // sp_3d_ns.cxx:772
    r->pp.obl     = 1;
    r->pp.obl_g   = r->deg;
    r->pp.obl_jump = r->sre * r->pp.L[1];
    // the sheared runs' LJ: every pair once (md_force.cxx:198-305), and the
    // diameter min((SIGMA + XI)/2^(1/2n), SIGMA) (input.cxx:549-566)
    r->pp.lj_cell_compat = 0;
    const double sigma   = 2.0 * r->radius;
    const double pw      = r->P.lj_powers == 0 ? pow(2., 1. / 6.) : r->P.lj_powers == 1 ? pow(2., 1. / 12.)
                                                                                       : pow(2., 1. / 18.);
    const double dia     = std::min((sigma + r->xi) / pw, sigma);
    r->pp.lj_cutoff      = sigma > 0.0 ? r->lj_cutoff / sigma * dia : 0.0;
    r->pp.lj_dia         = dia;
    return KH_OK;
}

int kh_get_shear_record(kh_run *r, double out[KH_SHEAR_RECORD]) {
    if (!r || !out) return fail(KH_ERR_ARG, "NULL argument");
    if (!r->shear) return fail(KH_ERR_STATE, "not a sheared run (kh_set_shear)");
    KH_HIP(hipMemcpyAsync(out, r->shear_last, sizeof(double) * KH_SHEAR_RECORD, hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}

int kh_set_shear_history(kh_run *r, int nsteps) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (!r->shear) return fail(KH_ERR_STATE, "not a sheared run (kh_set_shear)");
    if (nsteps < 0) return fail(KH_ERR_ARG, "negative step count");
    if (r->shear_hist) {
        KH_HIP(hipFreeAsync(r->shear_hist, r->stream));
        r->shear_hist = nullptr;
    }
    r->shear_hist_n = 0;
    if (nsteps == 0) return KH_OK;
    const size_t bytes = sizeof(double) * KH_SHEAR_RECORD * (size_t)nsteps;
    if (hipMallocAsync((void **)&r->shear_hist, bytes, r->stream) != hipSuccess) {
        r->shear_hist = nullptr;
        return fail(KH_ERR_NOMEM, "stress history of %zu bytes", bytes);
    }
    KH_HIP(hipMemsetAsync(r->shear_hist, 0, bytes, r->stream));
    r->shear_hist_n    = nsteps;
    r->shear_hist_base = r->jikan.ts;
    return KH_OK;
}

int kh_get_shear_history(kh_run *r, double *out, int nsteps) {
    if (!r || (!out && nsteps > 0)) return fail(KH_ERR_ARG, "NULL argument");
    if (nsteps < 0 || nsteps > r->shear_hist_n) return fail(KH_ERR_ARG, "%d records requested, %d kept", nsteps,
                                                            r->shear_hist_n);
    if (nsteps > 0)
        KH_HIP(hipMemcpyAsync(out, r->shear_hist, sizeof(double) * KH_SHEAR_RECORD * (size_t)nsteps,
                              hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}

int kh_get_shear_state(kh_run *r, double *degree_oblique, double *shear_rate_eff) {
    if (!r || !degree_oblique || !shear_rate_eff) return fail(KH_ERR_ARG, "NULL argument");
    *degree_oblique = r->deg;
    *shear_rate_eff = r->sre;
    return KH_OK;
}

int kh_set_shear_state(kh_run *r, double degree_oblique, double shear_rate_eff) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (!r->shear) return fail(KH_ERR_STATE, "kh_set_shear first");
    if (!std::isfinite(degree_oblique) || !std::isfinite(shear_rate_eff) || fabs(degree_oblique) >= r->xyaspect)
        return fail(KH_ERR_ARG, "strain must be finite and inside (-NX/NY, NX/NY)");
    KH_TRY(settle_zeta(r));  // at the strain the last step left
    r->deg         = degree_oblique;
    r->sre         = shear_rate_eff;
    r->pp.obl_g    = r->deg;
    r->pp.obl_jump = r->sre * r->pp.L[1];
    return KH_OK;
}

int kh_set_trajectory(kh_run *r, int nsteps) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (nsteps < 0) return fail(KH_ERR_ARG, "negative step count");
    drop_graph(r);  // the record kernel joins (or leaves) the captured step
    if (r->traj) {
        KH_HIP(hipFreeAsync(r->traj, r->stream));
        r->traj = nullptr;
    }
    r->traj_steps = 0;
    if (nsteps == 0 || r->np == 0) return KH_OK;
    const size_t bytes = sizeof(double) * 6 * (size_t)nsteps * (size_t)r->np;
    if (hipMallocAsync((void **)&r->traj, bytes, r->stream) != hipSuccess) {
        r->traj = nullptr;
        return fail(KH_ERR_NOMEM, "trajectory buffer of %zu bytes", bytes);
    }
    r->traj_steps = nsteps;
    r->traj_base  = r->jikan.ts;
    return KH_OK;
}

int kh_get_trajectory(kh_run *r, double *out, int nsteps) {
    if (!r || (!out && nsteps > 0)) return fail(KH_ERR_ARG, "NULL argument");
    if (nsteps < 0 || nsteps > r->traj_steps) return fail(KH_ERR_ARG, "%d steps requested, %d recorded", nsteps, r->traj_steps);
    if (nsteps == 0 || r->np == 0) return KH_OK;
    KH_HIP(hipMemcpyAsync(out, r->traj, sizeof(double) * 6 * (size_t)nsteps * (size_t)r->np, hipMemcpyDeviceToHost,
                          r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}

static_assert(KH_SHEAR_RECORD == 12, "k_obl_record writes 12 doubles");
static_assert(KH_FRAME_WORDS == FRAME_WORDS, "Output_udf frame layout");

int kh_set_output(kh_run *r, int nframes, int every) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (nframes < 0 || every < 1) return fail(KH_ERR_ARG, "nframes %d, every %d", nframes, every);
    drop_graph(r);  // the frame kernel joins (or leaves) the captured step
    if (r->frames) {
        KH_HIP(hipFreeAsync(r->frames, r->stream));
        r->frames = nullptr;
    }
    r->frame_n = 0;
    if (nframes == 0) return KH_OK;
    const size_t bytes = sizeof(double) * (size_t)nframes * (1 + (size_t)KH_FRAME_WORDS * (size_t)r->np);
    if (hipMallocAsync((void **)&r->frames, bytes, r->stream) != hipSuccess) {
        r->frames = nullptr;
        return fail(KH_ERR_NOMEM, "output buffer of %zu bytes", bytes);
    }
    // first frame at the first ts >= now with ts % every == 0 (the
    // reference writes before the step of that ts, sp_3d_ns.cxx:863-950)
    const int ts   = r->jikan.ts;
    r->frame_ts0   = (ts + every - 1) / every * every;
    r->frame_every = every;
    r->frame_n     = nframes;
    if (r->frame_ts0 == ts)  // the current state is frame 0
        KH_LAUNCH("frame", frame(r->pp, r->dp, r->d_ts, r->frame_ts0, every, nframes, r->frames, r->stream));
    return KH_OK;
}

int kh_get_output(kh_run *r, double *out, int nframes) {
    if (!r || (!out && nframes > 0)) return fail(KH_ERR_ARG, "NULL argument");
    if (nframes < 0 || nframes > r->frame_n) return fail(KH_ERR_ARG, "%d frames requested, %d set", nframes, r->frame_n);
    if (nframes == 0) return KH_OK;
    KH_HIP(hipMemcpyAsync(out, r->frames, sizeof(double) * (size_t)nframes * (1 + (size_t)KH_FRAME_WORDS * r->np),
                          hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}

int kh_set_rigid(kh_run *r, int nb, const int *cum, const int *free_flags, const double *v0, const double *w0) {
    using namespace rigid_init;
    if (!r || !cum) return fail(KH_ERR_ARG, "NULL argument");
    if (nb < 1 || cum[0] != 0 || cum[nb] != r->np) return fail(KH_ERR_ARG, "bead offsets must run from 0 to np");
    for (int k = 0; k < nb; k++)
        if (cum[k + 1] <= cum[k]) return fail(KH_ERR_ARG, "body %d has no beads", k);
    if (r->kbt > 0. || r->slip || r->shear || !r->P.rotation)
        return fail(KH_ERR_UNSUPPORTED, "rigid bodies need ROTATION on, kT = 0, no slip species and no shear");
    std::vector<DevParticle> h((size_t)r->np);
    KH_HIP(hipMemcpyAsync(h.data(), r->dp, sizeof(DevParticle) * h.size(), hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    const Geo g{{r->g.nx, r->g.ny, r->g.nz}, r->P.dx, {r->pp.L[0], r->pp.L[1], r->pp.L[2]}, r->radius, r->hxi};
    std::vector<DevBody> B((size_t)nb);
    std::vector<int>     bid((size_t)r->np);
    std::vector<double>  GR(3 * (size_t)r->np), GRb(3 * (size_t)r->np);
    for (int k = 0; k < nb; k++) {
        DevBody &b = B[k];
        memset(&b, 0, sizeof b);
        b.first = cum[k];
        b.last  = cum[k + 1];
        b.spec  = h[cum[k]].spec;
        b.free_ = free_flags ? (free_flags[k] != 0) : 1;
        // init_set_xGs (rigid.h:25-38): the trial centre, from the
        // unwrapped bead coordinates; init_set_PBC (58-80)
        for (int n = b.first; n < b.last; n++) {
            bid[n] = k;
            for (int d = 0; d < 3; d++) b.xG[d] += h[n].x_nopbc[d];
        }
        for (int d = 0; d < 3; d++) {
            b.xG[d] /= (double)(b.last - b.first);
            b.xGn[d] = b.xGp[d] = b.xG[d];
            b.xG[d]             = fmod(b.xG[d] + g.L[d], g.L[d]);
        }
        for (int n = b.first; n < b.last; n++)
            for (int d = 0; d < 3; d++) h[n].x_prev[d] = h[n].x_nopbc[d];
    }
    auto set_gr = [&]() {  // init_set_GRvecs (rigid.h:44-54)
        for (int n = 0; n < r->np; n++)
            for (int d = 0; d < 3; d++) GR[3 * n + d] = h[n].x_nopbc[d] - B[bid[n]].xGn[d];
    };
    set_gr();
    // Make_phi_particle_sum (make_phi.cxx:214-260) over every bead
    std::unordered_map<long long, double> phisum;
    for (int n = 0; n < r->np; n++)
        for (int code : r->sek_host) {
            const Pt p = point_h(g, h[n].x, code);
            phisum[p.idx] += phi_h(g, dist_h(g, p.x, h[n].x));
        }
    // phi/max(phi_sum,1)-weighted mass, first moment and inertia of a body
    // about its current centre (make_phi.cxx:589-630, 674-738)
    auto moments = [&](int k, double &m, double *c, double J[3][3]) {
        m = c[0] = c[1] = c[2] = 0.0;
        for (int d = 0; d < 3; d++) J[d][0] = J[d][1] = J[d][2] = 0.0;
        for (int n = B[k].first; n < B[k].last; n++)
            for (int code : r->sek_host) {
                const Pt     p = point_h(g, h[n].x, code);
                const double w = phi_h(g, dist_h(g, p.x, h[n].x)) / std::max(phisum[p.idx], 1.0);
                const double R[3] = {GR[3 * n] + p.r[0], GR[3 * n + 1] + p.r[1], GR[3 * n + 2] + p.r[2]};
                m += w;
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
    };
    for (int k = 0; k < nb; k++) {  // Make_phi_rigid_mass: mass, centre
        double m, c[3], J[3][3];
        moments(k, m, c, J);
        B[k].M = m * r->dx3 * r->species[B[k].spec].mass_ratio * r->rho;
        for (int d = 0; d < 3; d++) {
            B[k].xGn[d] += c[d] / m;
            B[k].xG[d] = fmod(B[k].xGn[d] + g.L[d], g.L[d]);
        }
    }
    set_gr();
    for (int k = 0; k < nb; k++) {  // Make_phi_rigid_inertia, init_Rigid_Coordinates (rigid.h:121-156)
        double m, c[3], J[3][3], V[3][3], ev[3], R[3][3], q[4];
        moments(k, m, c, J);
        const double rho_p = r->species[B[k].spec].mass_ratio * r->rho;
        for (int d = 0; d < 3; d++)
            for (int e = 0; e < 3; e++) J[d][e] = J[d][e] * r->dx3 * rho_p;
        if (!jacobi(J, V, ev)) return fail(KH_ERR_ARG, "body %d: inertia diagonalisation did not converge", k);
        frame(V);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[j][i] = V[i][j];
        to_quaternion(R, q);
        // principal moments: the diagonal of R^T J R
        for (int d = 0; d < 3; d++) {
            double col[3] = {J[0][d], J[1][d], J[2][d]}, rc[3];
            rotate(rc, q, col, true);
            for (int e = 0; e < 3; e++) R[e][d] = rc[e];
        }
        for (int d = 0; d < 3; d++) {
            double row[3] = {R[d][0], R[d][1], R[d][2]}, rr[3];
            rotate(rr, q, row, true);
            B[k].Ib[d] = rr[d];
        }
        for (int n = B[k].first; n < B[k].last; n++) {
            for (int a = 0; a < 4; a++) h[n].q[a] = q[a];
            rotate(&GRb[3 * n], q, &GR[3 * n], true);
        }
        // init_set_vGs (rigid.h:230-240): initial body velocities (space frame)
        for (int d = 0; d < 3; d++) {
            B[k].vG[d] = v0 ? v0[3 * k + d] : 0.0;
            B[k].wG[d] = w0 ? w0[3 * k + d] : 0.0;
        }
    }
    for (int n = 0; n < r->np; n++) {  // set_Particle_Velocities (rigid.h:218-229)
        const DevBody &b = B[bid[n]];
        double         wxr[3];
        cross(wxr, b.wG, &GR[3 * n]);
        for (int d = 0; d < 3; d++) {
            h[n].om[d] = b.wG[d];
            h[n].v[d]  = b.vG[d] + wxr[d];
        }
    }
    // device side
    drop_graph(r);
    for (void *b : {(void *)r->ra.B, (void *)r->ra.GR, (void *)r->ra.GRb, (void *)r->bid, (void *)r->flo})
        if (b) KH_HIP(hipFreeAsync(b, r->stream));
    r->ra    = RigidArgs{};
    r->bid   = nullptr;
    r->flo   = nullptr;
    r->rigid = 0;
    auto alloc = [&](void **p, size_t bytes) -> int {
        if (hipMallocAsync(p, bytes, r->stream) != hipSuccess) {
            *p = nullptr;
            return fail(KH_ERR_NOMEM, "rigid body buffers");
        }
        return KH_OK;
    };
    KH_TRY(alloc((void **)&r->ra.B, sizeof(DevBody) * (size_t)nb));
    KH_TRY(alloc((void **)&r->ra.GR, sizeof(double) * GR.size()));
    KH_TRY(alloc((void **)&r->ra.GRb, sizeof(double) * GRb.size()));
    KH_TRY(alloc((void **)&r->bid, sizeof(int) * bid.size()));
    KH_TRY(alloc((void **)&r->flo, sizeof(double) * 3 * (size_t)r->np));
    KH_HIP(hipMemcpyAsync(r->ra.B, B.data(), sizeof(DevBody) * B.size(), hipMemcpyHostToDevice, r->stream));
    KH_HIP(hipMemcpyAsync(r->ra.GR, GR.data(), sizeof(double) * GR.size(), hipMemcpyHostToDevice, r->stream));
    KH_HIP(hipMemcpyAsync((void *)r->ra.GRb, GRb.data(), sizeof(double) * GRb.size(), hipMemcpyHostToDevice,
                          r->stream));
    KH_HIP(hipMemcpyAsync(r->bid, bid.data(), sizeof(int) * bid.size(), hipMemcpyHostToDevice, r->stream));
    KH_HIP(hipMemsetAsync(r->flo, 0, sizeof(double) * 3 * (size_t)r->np, r->stream));
    KH_HIP(hipMemcpyAsync(r->dp, h.data(), sizeof(DevParticle) * h.size(), hipMemcpyHostToDevice, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    r->ra.nb      = nb;
    r->ra.compat  = r->P.lj_cell_compat;
    r->pp.bid     = r->bid;
    r->pp.flo     = r->flo;
    r->rigid      = 1;
    r->pcache_valid = 0;
    return KH_OK;
}

int kh_get_rigid(kh_run *r, double *out, int nb) {
    if (!r || (!out && nb > 0)) return fail(KH_ERR_ARG, "NULL argument");
    if (!r->rigid || nb != r->ra.nb) return fail(KH_ERR_ARG, "%d bodies requested, the run has %d", nb, r->ra.nb);
    std::vector<DevBody>     B((size_t)nb);
    std::vector<DevParticle> h((size_t)r->np);
    KH_HIP(hipMemcpyAsync(B.data(), r->ra.B, sizeof(DevBody) * B.size(), hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipMemcpyAsync(h.data(), r->dp, sizeof(DevParticle) * h.size(), hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    for (int k = 0; k < nb; k++) {
        const DevBody &b = B[k];
        double        *o = out + (size_t)KH_RIGID_WORDS * k;
        for (int d = 0; d < 3; d++) {
            o[d]      = b.xG[d];
            o[3 + d]  = b.xGn[d];
            o[6 + d]  = b.vG[d];
            o[9 + d]  = b.wG[d];
            o[12 + d] = b.fGp[d];
            o[15 + d] = b.tGp[d];
            o[18 + d] = b.fGrp[d];
            o[21 + d] = b.tGrp[d];
            o[25 + d] = b.Ib[d];
        }
        o[24] = b.M;
        for (int a = 0; a < 4; a++) o[28 + a] = h[b.first].q[a];
    }
    return KH_OK;
}

long long kh_run_bytes(const kh_params *P, int np) {
    if (!P || np < 0) return -1;
    return (long long)run_footprint(*P, np);
}

int kh_query(kh_run *r) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    const hipError_t e = hipStreamQuery(r->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(KH_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int kh_set_graph(kh_run *r, int on) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    r->use_graph = on != 0;
    return KH_OK;
}

int kh_apply(kh_run *r, int fn) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    KH_TRY(settle_zeta(r));
    return r->mixed ? Step<float>::apply(r, fn) : Step<double>::apply(r, fn);
}

int kh_sync(kh_run *r) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    KH_TRY(settle_zeta(r));
    KH_HIP(hipStreamSynchronize(r->stream));
    return KH_OK;
}

int kh_A2a_k(kh_run *r, double *a) {
    if (!r || !a) return fail(KH_ERR_ARG, "NULL argument");
    KH_TRY(upload_real(r, r->work[4], a));
    if (r->mixed) {
        KH_TRY(Step<float>::r2c(r, (float *)r->work[4], (cpx<float> *)r->spec[4]));
    } else {
        KH_TRY(Step<double>::r2c(r, (double *)r->work[4], (cpx<double> *)r->spec[4]));
    }
    return download_spec(r, a, r->spec[4]);
}

int kh_A_k2a(kh_run *r, double *a) {
    if (!r || !a) return fail(KH_ERR_ARG, "NULL argument");
    KH_TRY(upload_spec(r, r->spec[4], a));
    if (r->mixed) {
        KH_TRY(Step<float>::c2r(r, (cpx<float> *)r->spec[4], (float *)r->work[4]));
    } else {
        KH_TRY(Step<double>::c2r(r, (cpx<double> *)r->spec[4], (double *)r->work[4]));
    }
    return download_real(r, a, r->work[4]);
}

int kh_get_slip_iterations(kh_run *r, long long *iterations, long long *loops) {
    if (!r || !iterations || !loops) return fail(KH_ERR_ARG, "NULL argument");
    unsigned long long st[2] = {0, 0};
    KH_HIP(hipMemcpyAsync(st, (char *)r->slipflag + 16, sizeof st, hipMemcpyDeviceToHost, r->stream));
    KH_HIP(hipStreamSynchronize(r->stream));
    *iterations = (long long)st[0] + r->slip_iters_host;
    *loops      = (long long)st[1] + r->slip_loops_host;
    return KH_OK;
}

int kh_set_step_path(kh_run *r, int path) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    if (path != KH_PATH_FUSED && path != KH_PATH_REFERENCE_ORDER) return fail(KH_ERR_ARG, "unknown step path %d", path);
    KH_TRY(settle_zeta(r));
    r->step_path = path;
    drop_graph(r);
    return KH_OK;
}

int kh_timing_enable(kh_run *r, int on) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    r->timing = on != 0;
    return KH_OK;
}
int kh_timing_reset(kh_run *r) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    KH_TRY(resolve_timing(r));
    r->times.clear();
    return KH_OK;
}
int kh_timing_get(kh_run *r, const char *name, int *launches, double *total_ms) {
    if (!r || !name) return fail(KH_ERR_ARG, "NULL argument");
    KH_TRY(resolve_timing(r));
    auto it = r->times.find(name);
    if (launches) *launches = it == r->times.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == r->times.end() ? 0.0 : it->second.ms;
    return KH_OK;
}
int kh_timing_count(kh_run *r) {
    if (!r) return fail(KH_ERR_ARG, "NULL run");
    KH_TRY(resolve_timing(r));
    r->names.clear();
    for (auto &kv : r->times) r->names.push_back(kv.first);
    return (int)r->names.size();
}
const char *kh_timing_name(kh_run *r, int i) {
    if (!r || i < 0 || i >= (int)r->names.size()) return nullptr;
    return r->names[i].c_str();
}


// ------------------------------------------------- allocation-free transform
// kh_A2a_k_noalloc: the forward transform of kh_A2a_k without a run and
// without any device allocation. Its staging (the real field and its
// spectrum) lives in `scratch`, a file-scope device array of fixed size that
// each GPU the library runs on holds until the process ends. Host code never
// passes the array's name to a kernel or a copy (on the host it is only a
// stand-in): the device address is taken once per device with
// hipGetSymbolAddress and kept in noalloc_ptr. Every call on one GPU uses
// the same array, so calls that overlap on one GPU race.
}  // extern "C"

namespace {
constexpr size_t KH_NOALLOC_MAX = KH_NOALLOC_MAX_BYTES;  // bytes of scratch per device
constexpr int    KH_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[KH_NOALLOC_MAX];
std::mutex noalloc_mu;
void      *noalloc_ptr[KH_MAX_DEVICES] = {};

// the device address of `scratch` on the current device (taken on the first
// call there)
int noalloc_scratch(void **out) {
    int dev = 0;
    KH_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= KH_MAX_DEVICES) return fail(KH_ERR_UNSUPPORTED, "device %d beyond %d", dev, KH_MAX_DEVICES);
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        KH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(KH_ERR_HIP, "hipGetSymbolAddress returned a null address");
        noalloc_ptr[dev] = p;
    }
    *out = noalloc_ptr[dev];
    return KH_OK;
}
}  // namespace

extern "C" {

int kh_A2a_k_noalloc(double *a, int nx, int ny, int nz) {
    if (!a) return fail(KH_ERR_ARG, "NULL argument");
    if (nx < 2 || ny < 2 || nz < 2 || (nz & 1))
        return fail(KH_ERR_ARG, "grid %d x %d x %d: sizes must be at least 2, NZ even", nx, ny, nz);
    if (!fft_length_supported(nx) || !fft_length_supported(ny) || !fft_length_supported(nz))
        return fail(KH_ERR_UNSUPPORTED, "grid %d x %d x %d: lengths must be 2^a 3^b from 8 to 512", nx, ny, nz);
    Grid g{};
    g.nx  = nx;
    g.ny  = ny;
    g.nz  = nz;
    g.nzh = nz / 2 + 1;
    g.nzp = (g.nzh + 7) / 8 * 8;  // FP64 spectral rows on 128-byte boundaries
    g.hnx = nx / 2;
    g.hny = ny / 2;
    g.trn_x = (nx + 2) / 3;
    g.trn_y = (ny + 2) / 3;
    g.trn_z = (nz + 2) / 3;
    g.wave_x = 2. * KH_PI / nx;
    g.wave_y = 2. * KH_PI / ny;
    g.wave_z = 2. * KH_PI / nz;
    g.nreal  = (long long)nx * ny * nz;
    g.nspec  = (long long)nx * ny * g.nzp;
    const size_t real_bytes = ((size_t)g.nreal * sizeof(double) + 255) / 256 * 256;
    const size_t need       = real_bytes + (size_t)g.nspec * sizeof(cpx<double>);
    if (need > KH_NOALLOC_MAX)
        return fail(KH_ERR_UNSUPPORTED, "grid %d x %d x %d needs %zu bytes of scratch, more than the %zu available",
                    nx, ny, nz, need, KH_NOALLOC_MAX);
    void *base = nullptr;
    KH_TRY(noalloc_scratch(&base));
    double      *real = (double *)base;
    cpx<double> *spec = (cpx<double> *)((char *)base + real_bytes);
    const int    nz_  = 2 * g.nzh;
    hipStream_t  s    = nullptr;  // the null stream: nothing is created
    KH_HIP(hipMemcpy2DAsync(real, sizeof(double) * nz, a, sizeof(double) * nz_, sizeof(double) * nz,
                            (size_t)nx * ny, hipMemcpyHostToDevice, s));
    KH_HIP(fft3d_r2c<double>(g, real, spec, s));
    KH_HIP(hipMemcpy2DAsync(a, sizeof(double) * nz_, spec, sizeof(cpx<double>) * g.nzp, sizeof(double) * nz_,
                            (size_t)nx * ny, hipMemcpyDeviceToHost, s));
    KH_HIP(hipStreamSynchronize(s));
    return KH_OK;
}

}  // extern "C"

// ------------------------------------------------------------ batches
namespace kh {

size_t batch_stride(const kh_params &P, int np_cap) {
    // the footprint plus the 256-byte alignment of every buffer
    return (run_footprint(P, np_cap) + 65536 + 65535) / 65536 * 65536;
}

bool batch_accepts(const kh_batch *b, const kh_params &P, int np) {
    const kh_params &Q = b->P;
    if (np > b->np_cap) return false;
    if (P.nx != Q.nx || P.ny != Q.ny || P.nz != Q.nz || P.precision != Q.precision || P.dx != Q.dx ||
        P.rho != Q.rho || P.eta != Q.eta || P.dt != Q.dt || P.a != Q.a || P.a_xi != Q.a_xi)
        return false;
    for (int d = 0; d < 3; d++)
        if ((P.fix_cell[d] != 0) != (Q.fix_cell[d] != 0)) return false;
    for (int i = 0; i < P.nspecies; i++)
        if (P.species[i].propulsion == KH_JP_SLIP) return false;
    return run_footprint(P, np) <= b->stride;
}

int batch_create(const kh_params &P, int np_cap, int capacity, kh_batch **out) {
    if (!out || capacity < 1 || np_cap < 1 || np_cap > 256) return fail(KH_ERR_ARG, "bad batch arguments");
    *out        = nullptr;
    kh_batch *b = new kh_batch();
    b->P        = P;
    b->P.species  = nullptr;
    b->P.nspecies = 0;
    b->capacity = capacity;
    b->np_cap   = np_cap;
    b->stride   = batch_stride(P, np_cap);
    b->hactive.assign(capacity, 0);
    b->hslots.assign(sizeof(SlotPart<double>) * (size_t)capacity, 0);
    b->slot_run.assign(capacity, nullptr);
    auto bail = [&](int st) {
        if (b->stream) (void)hipStreamSynchronize(b->stream);
        if (b->base) (void)hipFree(b->base);
        if (b->active) (void)hipFree(b->active);
        if (b->dslots) (void)hipFree(b->dslots);
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return st;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(KH_ERR_HIP, "hipStreamCreate failed"));
    if (hipMallocAsync((void **)&b->base, b->stride * (size_t)capacity, b->stream) != hipSuccess) {
        b->base = nullptr;
        return bail(fail(KH_ERR_NOMEM, "batch arena of %zu bytes", b->stride * (size_t)capacity));
    }
    if (hipMallocAsync((void **)&b->active, sizeof(int) * (size_t)capacity, b->stream) != hipSuccess) {
        b->active = nullptr;
        return bail(fail(KH_ERR_NOMEM, "batch flags"));
    }
    if (hipMallocAsync(&b->dslots, b->hslots.size(), b->stream) != hipSuccess) {
        b->dslots = nullptr;
        return bail(fail(KH_ERR_NOMEM, "batch slot table"));
    }
    if (hipMemcpyAsync(b->active, b->hactive.data(), sizeof(int) * (size_t)capacity, hipMemcpyHostToDevice,
                       b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bail(fail(KH_ERR_HIP, "batch set-up failed"));
    *out = b;
    return KH_OK;
}

int batch_free_slot(const kh_batch *b) {
    for (int k = 0; k < b->capacity; k++)
        if (!b->slot_run[k]) return k;
    return -1;
}

int batch_run_create(kh_batch *b, int slot, const kh_params *P, int np, kh_run **out) {
    if (!b || !P || !out || slot < 0 || slot >= b->capacity || b->slot_run[slot])
        return fail(KH_ERR_ARG, "slot %d is not free", slot);
    if (!batch_accepts(b, *P, np)) return fail(KH_ERR_ARG, "run is not of the batch's class");
    KH_TRY(run_create(P, np, b->base + (size_t)slot * b->stride, b->stride, out));
    (*out)->slot     = slot;
    b->slot_run[slot] = *out;
    return KH_OK;
}

static void batch_drop_graphs(kh_batch *b) {
    if (b->exec1) (void)hipGraphExecDestroy(b->exec1);
    if (b->graph1) (void)hipGraphDestroy(b->graph1);
    if (b->execk) (void)hipGraphExecDestroy(b->execk);
    if (b->graphk) (void)hipGraphDestroy(b->graphk);
    b->exec1 = b->execk = nullptr;
    b->graph1 = b->graphk = nullptr;
}

int batch_join(kh_batch *b, kh_run *r) {
    if (!b || !r || r->slot < 0 || r->slot >= b->capacity || b->slot_run[r->slot] != r)
        return fail(KH_ERR_ARG, "run is not in this batch");
    if (r->slip || r->rigid || r->shear || r->step_path != KH_PATH_FUSED || r->jikan.ts < 1 || r->zeta_dirty ||
        r->timing)
        return fail(KH_ERR_STATE, "only plain runs past their first fused step join a batch");
    // the batched launches address every run's buffers from the first one's
    if (r->arena != b->base + (size_t)r->slot * b->stride || (char *)r->zeta[0] != r->arena)
        return fail(KH_ERR_STATE, "run's buffers are not at the batch's offsets");
    KH_HIP(hipStreamSynchronize(r->own_stream));
    const size_t esz = r->mixed ? sizeof(SlotPart<float>) : sizeof(SlotPart<double>);
    char        *h   = b->hslots.data() + (size_t)r->slot * esz;
    if (r->mixed) {
        const SlotPart<float> a = Step<float>::slot_part(r);
        memcpy(h, &a, esz);
    } else {
        const SlotPart<double> a = Step<double>::slot_part(r);
        memcpy(h, &a, esz);
    }
    KH_HIP(hipMemcpyAsync((char *)b->dslots + (size_t)r->slot * esz, h, esz, hipMemcpyHostToDevice, b->stream));
    r->stream  = b->stream;
    r->batched = true;
    b->hactive[r->slot] = 1;
    b->dirty            = true;
    return KH_OK;
}

int batch_leave(kh_batch *b, kh_run *r) {
    if (!b || !r || r->slot < 0 || r->slot >= b->capacity || b->slot_run[r->slot] != r)
        return fail(KH_ERR_ARG, "run is not in this batch");
    KH_HIP(hipStreamSynchronize(b->stream));
    r->stream  = r->own_stream;
    r->batched = false;
    b->hactive[r->slot]  = 0;
    b->slot_run[r->slot] = nullptr;
    b->dirty             = true;
    return KH_OK;
}

int batch_members(const kh_batch *b) {
    int n = 0;
    for (int k = 0; k < b->capacity; k++) n += b->hactive[k];
    return n;
}

// Captures `count` batch steps over every slot, through a view of slot 0
// made from the joined run m.
template <typename T>
static int batch_capture(kh_batch *b, kh_run *m, int count, hipGraph_t *graph, hipGraphExec_t *exec) {
    std::unique_ptr<kh_run> v(new kh_run(*m));
    const long long         off = -(long long)m->slot * (long long)b->stride;
    auto                    mv  = [&](auto *&p) {
        using P = std::remove_reference_t<decltype(p)>;
        p       = reinterpret_cast<P>(reinterpret_cast<char *>(p) + off);
    };
    for (auto &p : v->zeta) mv(p);
    for (auto &p : v->spec) mv(p);
    for (auto &p : v->ureal) mv(p);
    for (auto &p : v->up) mv(p);
    for (auto &p : v->work) mv(p);
    mv(v->phisum);
    mv(v->occ);
    mv(v->ukdc);
    mv(v->d_ts);
    v->slot      = 0;
    v->stream    = b->stream;
    v->timing    = false;
    v->capturing = true;
    v->pending.clear();
    v->event_pool.clear();
    KH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int st = KH_OK;
    for (int k = 0; k < count && st == KH_OK; k++) st = Step<T>::step_batch(b, v.get());
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (st != KH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(KH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    *graph = g;
    KH_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return KH_OK;
}

int batch_step(kh_batch *b, int nsteps) {
    if (!b || nsteps < 0) return fail(KH_ERR_ARG, "bad batch step");
    std::vector<kh_run *> members;
    for (int k = 0; k < b->capacity; k++)
        if (b->hactive[k]) members.push_back(b->slot_run[k]);
    if (b->dirty) {  // the slots that step, from the next launch on
        KH_HIP(hipMemcpyAsync(b->active, b->hactive.data(), sizeof(int) * (size_t)b->capacity, hipMemcpyHostToDevice,
                              b->stream));
        KH_HIP(hipStreamSynchronize(b->stream));
        b->dirty = false;
    }
    if (members.empty() || nsteps == 0) return KH_OK;
    const bool mixed = b->P.precision == KH_MIXED;
    for (int s = 0; s < nsteps;) {
        int done = 1;
        if (nsteps - s >= GRAPH_STEPS) {
            if (!b->execk)
                KH_TRY(mixed ? batch_capture<float>(b, members[0], GRAPH_STEPS, &b->graphk, &b->execk)
                             : batch_capture<double>(b, members[0], GRAPH_STEPS, &b->graphk, &b->execk));
            KH_HIP(hipGraphLaunch(b->execk, b->stream));
            done = GRAPH_STEPS;
        } else {
            if (!b->exec1)
                KH_TRY(mixed ? batch_capture<float>(b, members[0], 1, &b->graph1, &b->exec1)
                             : batch_capture<double>(b, members[0], 1, &b->graph1, &b->exec1));
            KH_HIP(hipGraphLaunch(b->exec1, b->stream));
        }
        for (kh_run *q : members) {
            q->pcache_valid = 1;
            q->pclean       = q->zclear;
            for (int k = 0; k < done; k++) {
                q->jikan.time += q->jikan.dt_fluid;  // sp_3d_ns.cxx:955
                q->jikan.ts++;
            }
        }
        s += done;
    }
    return KH_OK;
}

int batch_query(kh_batch *b) {
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return fail(KH_ERR_HIP, "hipStreamQuery: %s", hipGetErrorString(e));
}

int batch_destroy(kh_batch *b) {
    if (!b) return KH_OK;
    for (kh_run *q : b->slot_run)
        if (q) return fail(KH_ERR_STATE, "batch still holds a run");
    (void)hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    (void)hipFreeAsync(b->base, b->stream);
    (void)hipFreeAsync(b->active, b->stream);
    (void)hipFreeAsync(b->dslots, b->stream);
    (void)hipStreamSynchronize(b->stream);
    (void)hipStreamDestroy(b->stream);
    delete b;
    return KH_OK;
}

}  // namespace kh
