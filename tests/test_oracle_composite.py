"""Pins the oracle's composite step, Time_evolution_hydro
(sp_3d_ns.cxx:41-194), to what the mathematics fixes, independently of any
transcription of its code:

  * one step from a fluid at rest in closed form (Euler first step, F_H at
    the moved position with the old velocity, its -rho dx^3/dt prefactor,
    the rigid-body velocity v + omega x r, the projected u_p as the new
    fluid);
  * two steps in the linear regime in closed form (amplitudes of 1e-20,
    where advection is below round-off): the exact exp(-nu k^2 dt) decay of
    the first step's flow as u*, FIX_CELL on the mean, F_H from u* with the
    old velocity, the AB2 update with 2 F_H (particle_solver.cxx:239-281);
  * an exact per-step momentum balance of particles and fluid with
    overlapping interfaces, LJ contacts and rotation in the nonlinear
    regime: what the particles gain, dt sum F_H, and what the fluid gains,
    rho dx^3 sum_g f, differ by rho dx^3 sum_i (W_i dv_i + d omega_i x R_i),
    W_i and R_i the brute-force weights sum_g phi_i / max(phi_sum, 1) and
    their first moments at the new positions;
  * exact Galilean invariance of a uniform flow carrying its particles;
  * a squirmer's swimming speed 2 B1 / 3 and a rotlet's rotation -C1 / a
    (reciprocal theorem, Stone & Samuel 1996), and zero rotation by symmetry.

Then each plausible mistake is planted in the oracle (kref.mutation) and at
least one of these checks must fail. CPU only; the brute-force sums use
the profile phi (profile.h:30-50) pinned in test_oracle_pins.py.
"""
import numpy as np
import pytest

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import JAX_Z, JP_NONE, JP_SLIP, Params, make_species, new_particles

N = 32
A = 4.0
XI = 2.0
RHO_P = 1.5  # particle / fluid density: M differs from the fluid mass it displaces


def phi_np(r, a=A, xi=XI, dx=1.0):
    """Phi (profile.h:30-50) on an array of distances."""
    def H(x):
        out = np.zeros_like(x)
        m = x > 0
        out[m] = np.exp(-(dx / x[m]) ** 2)
        return out
    h1 = H(a + 0.5 * xi - r)
    h2 = H(r - a + 0.5 * xi)
    s = h1 + h2
    out = np.zeros_like(r)
    m = s > 0
    out[m] = h1[m] / s[m]
    return out


def test_phi_np_matches_pinned_oracle_profile():
    rs = np.linspace(0.0, 7.0, 57)
    assert np.abs(phi_np(rs) - np.array([kref.phi(x, A, 1.0, XI) for x in rs])).max() < 1e-15


GRID = np.stack(np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij"), -1).astype(float)


def profile(x):
    """phi of a sphere at x on every grid point, and the minimum-image
    vectors r = g - x (the reference's relative coordinate while the
    Sekibun cell is smaller than half the box)."""
    r = GRID - np.asarray(x)
    r -= np.rint(r / N) * N
    return phi_np(np.sqrt((r ** 2).sum(-1))), r


def kgrid():
    k = np.fft.fftfreq(N, 1.0 / N)
    kz = np.arange(N // 2 + 1)
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    trn = (N + 2) // 3
    mask = (np.abs(KX) < trn) & (np.abs(KY) < trn) & (KZ < trn)  # 2/3 rule (fft_wrapper.h:754-796)
    w = 2 * np.pi / N
    return np.stack([KX * w, KY * w, KZ * w]), mask


def project(f):
    """IFFT(mask P FFT(f)) per component: the solenoidal, dealiased field
    (Solenoidal_u, operate_omega.h:105-112), P = I - k k / k^2 (identity at k = 0)."""
    K, mask = kgrid()
    F = np.stack([np.fft.rfftn(f[..., d]) for d in range(3)])
    k2 = (K ** 2).sum(0)
    k2[0, 0, 0] = 1.0
    kf = (K * F).sum(0)
    F = (F - K * kf / k2) * mask
    return np.stack([np.fft.irfftn(F[d], s=(N, N, N), axes=(0, 1, 2)) for d in range(3)], -1)


def decay(u, nu, dt, fix):
    """exp(-nu k^2 dt) on every mode of a band-limited field; the k = 0 mode
    kept unless FIX_CELL zeroes it."""
    K, mask = kgrid()
    k2 = (K ** 2).sum(0)
    out = []
    for d in range(3):
        F = np.fft.rfftn(u[..., d]) * np.exp(-nu * k2 * dt) * mask
        if fix[d]:
            F[0, 0, 0] = 0.0
        out.append(np.fft.irfftn(F, s=(N, N, N), axes=(0, 1, 2)))
    return np.stack(out, -1)


def params(fix=(0, 0, 0), rotation=1, lj=-1, **kw):
    return Params(N, a=A, a_xi=XI, rho=1.0, eta=1.0, dt=0.05, kbt=0.0, g=0.0, rotation=rotation,
                  fix_cell=fix, lj_truncate=lj, lj_cell_compat=0, species=[make_species(mass_ratio=RHO_P)], **kw)


def fluid_velocity(r):
    """u from the oracle's zeta and uk_dc (Zeta_k2u, itself pinned by the
    single-mode curl tests), as an (N, N, N, 3) array."""
    r.apply("zeta_k2u")
    return np.stack([r.get_field(f"u{d}")[:, :, :N] for d in range(3)], -1)


def start(P, x, v, om):
    r = kref.Run(P, len(x))
    p = new_particles(len(x))
    p["x"] = p["x_nopbc"] = p["x_previous"] = x
    p["v"] = v
    p["omega"] = om
    r.set_particles(p)
    z = np.zeros(r.shape)
    r.set_fluid(z, z, np.zeros(3))
    return r


MASS = RHO_P * 4.0 / 3.0 * np.pi * A ** 3
MOI = 2.0 / 5.0 * MASS * A ** 2
DT = 0.05


def rigid_impulse(x, v, om, u=None):
    """-(rho dx^3/dt) sum_g phi (v + omega x r - u) and its moment r x (.)
    for one sphere (Calc_f_hydro_correct_precision, md_force.cxx:383-506,
    with phi / max(phi_sum, 1) = phi for a lone sphere)."""
    ph, r = profile(x)
    V = v + np.cross(om, r)
    if u is not None:
        V = V - u
    F = -(1.0 / DT) * (ph[..., None] * V).sum((0, 1, 2))
    T = -(1.0 / DT) * (ph[..., None] * np.cross(r, V)).sum((0, 1, 2))
    return F, T


def one_step_closed_form(x0, v0, om0):
    x1 = np.mod(x0 + DT * v0, N)                  # Euler (particle_solver.cxx:39-61)
    F, T = rigid_impulse(x1, v0, om0)             # u* = 0: the fluid is at rest
    v1 = v0 + DT / MASS * F                        # MD_solver_velocity_Euler (103-142)
    om1 = om0 + DT / MOI * T
    ph, r = profile(x1)
    up = ph[..., None] * (v1 + np.cross(om1, r))   # Make_u_particle_sum
    u1 = project(up)                               # u = u* + P(u_p - phi u*), u* = 0
    return x1, v1, om1, u1, up.sum((0, 1, 2))


X0 = np.array([15.3, 16.6, 14.2])
V0 = np.array([0.05, -0.08, 0.11])
OM0 = np.array([0.02, -0.03, 0.015])


def check_one_step():
    r = start(params(), X0[None], V0[None], OM0[None])
    r.step(1)
    p = r.get_particles()[0]
    x1, v1, om1, u1, ukdc = one_step_closed_form(X0, V0, OM0)
    _, _, uk = r.get_fluid()
    u = fluid_velocity(r)
    errs = [np.abs(p["x"] - x1).max(),
            np.abs(p["v"] - v1).max() / np.abs(v1).max(),
            np.abs(p["omega"] - om1).max() / np.abs(om1).max(),
            np.abs(uk - ukdc).max() / np.abs(ukdc).max(),
            np.abs(u - u1).max() / np.abs(u1).max()]
    return max(errs), errs


def test_one_step_from_rest_closed_form():
    err, errs = check_one_step()
    assert err < 1e-12, errs


S = 1e-20  # linear regime: advection (S^2) is 1e-20 of the linear terms


def check_two_steps(fix):
    P = params(fix=fix)
    v0, om0 = S * V0, S * OM0
    r = start(P, X0[None], v0[None], om0[None])
    r.step(2)
    p = r.get_particles()[0]
    _, _, uk = r.get_fluid()
    u = fluid_velocity(r)
    # step 1 (ts = 0): as from rest
    x1, v1, om1, u1, _ = one_step_closed_form(X0, v0, om0)
    # step 2 (ts = 1): NS on u1 is the pure viscous decay; FIX_CELL zeroes the mean
    ustar = decay(u1, 1.0, DT, fix)
    x2 = np.mod(x1 + 0.5 * DT * (3 * v1 - v0), N)  # MD_solver_position_AB2
    F, T = rigid_impulse(x2, v1, om1, ustar)      # F_H at x2 from u* with the old v
    v2 = v1 + 0.5 * DT / MASS * (2.0 * F)          # MD_solver_velocity_AB2_hydro: 2 F_H
    om2 = om1 + 0.5 * DT / MOI * (2.0 * T)
    ph, rr = profile(x2)
    f = ph[..., None] * (v2 + np.cross(om2, rr)) - ph[..., None] * ustar
    u2 = ustar + project(f)
    errs = [np.abs(p["v"] - v2).max() / np.abs(v2).max(),
            np.abs(p["omega"] - om2).max() / np.abs(om2).max(),
            np.abs(u - u2).max() / np.abs(u2).max(),
            np.abs(uk - u2.sum((0, 1, 2))).max() / np.abs(u2.sum((0, 1, 2))).max()]
    return max(errs), errs


@pytest.mark.parametrize("fix", [(0, 0, 0), (1, 0, 1)])
def test_two_steps_linear_regime_closed_form(fix):
    err, errs = check_two_steps(fix)
    assert err < 1e-10, errs


# three spheres: interfaces of the first two overlap (phi_sum > 1 between
# them), the last two are inside the LJ cutoff 2^(1/6) 2a
XM = np.array([[10.2, 16.3, 15.7], [18.1, 16.6, 16.4], [18.4, 24.9, 16.2]])
VM = np.array([[0.012, -0.004, 0.006], [-0.009, 0.007, 0.003], [0.002, -0.011, -0.008]])
OMM = np.array([[0.003, -0.002, 0.004], [-0.001, 0.005, 0.002], [0.004, 0.001, -0.003]])


def weights(xs):
    """W_i = sum_g phi_i / max(phi_sum, 1) and R_i = sum_g (same) r_i, brute force."""
    phs, rs = zip(*[profile(x) for x in xs])
    tot = np.maximum(sum(phs), 1.0)
    W = np.array([(ph / tot).sum() for ph in phs])
    R = np.array([((ph / tot)[..., None] * r).sum((0, 1, 2)) for ph, r in zip(phs, rs)])
    return W, R


def check_momentum_balance(steps=8):
    P = params(lj=1)
    r = kref.Run(P, 3)
    p = new_particles(3)
    p["x"] = p["x_nopbc"] = p["x_previous"] = XM
    p["v"] = VM
    p["omega"] = OMM
    r.set_particles(p)
    r.init_fluid()
    worst = 0.0
    for _ in range(steps):
        a = r.get_particles()
        _, _, uka = r.get_fluid()
        r.step(1)
        b = r.get_particles()
        _, _, ukb = r.get_fluid()
        W, R = weights(b["x"])
        dv, dom = b["v"] - a["v"], b["omega"] - a["omega"]
        lhs = MASS * dv.sum(0) + (ukb - uka)
        rhs = (W[:, None] * dv).sum(0) + np.cross(dom, R).sum(0)
        worst = max(worst, np.abs(lhs - rhs).max() / (MASS * np.abs(dv).max()))
    return worst


def test_momentum_balance_overlapping_spheres_with_contacts():
    assert check_momentum_balance() < 1e-11


UG = np.array([0.013, -0.007, 0.021])
XG = np.array([[8.3, 9.1, 20.7], [22.4, 19.8, 7.6]])


def check_galilean(steps=6):
    P = params()
    r = kref.Run(P, 2)
    p = new_particles(2)
    p["x"] = p["x_nopbc"] = p["x_previous"] = XG
    p["v"] = UG
    r.set_particles(p)
    z = np.zeros(r.shape)
    r.set_fluid(z, z, UG * N ** 3)  # uk_dc: the unnormalised k = 0 mode, N^3 <u>
    r.step(steps)
    q = r.get_particles()
    z0, z1, uk = r.get_fluid()
    errs = [np.abs(q["v"] - UG).max() / np.abs(UG).max(),
            np.abs(q["omega"]).max() / np.abs(UG).max(),
            np.abs(q["x_nopbc"] - (XG + steps * DT * UG)).max(),
            np.abs(uk / N ** 3 - UG).max() / np.abs(UG).max(),
            max(np.abs(z0).max(), np.abs(z1).max()) / (np.abs(UG).max() * N ** 3)]
    return max(errs), errs


def test_galilean_invariance_of_a_uniform_flow():
    err, errs = check_galilean()
    assert err < 1e-13, errs


# ------------------------------------------------------------- squirmers
NS = 32
B1 = 0.01
C1 = 0.01


def squirmer_run(mode, steps=400, x=(16.0, 16.0, 16.0)):
    """One squirmer a = 4 along +z in 32^3 (volume fraction 0.008) with
    slip B1 (sin theta + (mode / 2) sin 2 theta) theta_hat + C1 sin theta
    phi_hat, from rest; slip tolerance 1e-4, up to 8 iterations. Returns the
    particle and the fluid velocity."""
    sp = make_species(mass_ratio=1.0, propulsion=JP_SLIP, janus_axis=JAX_Z, slip_vel=B1, slip_mode=mode,
                      rotlet_c1=C1)
    P = Params(NS, a=A, a_xi=XI, dt=0.05, rotation=1, fix_cell=(0, 0, 0), lj_truncate=-1, species=[sp],
               slip_iter_max=8, slip_tol=1e-4)
    r = kref.Run(P, 1)
    r.set_particles(workloads.particles_at(np.array([x])))
    r.init_fluid()
    r.step(steps)
    return r.get_particles()[0], fluid_velocity(r)


def check_squirmer_physics():
    """Two squirmers, B2 = +B1 and -B1 (mode +-1: the input B2/B1 is halved
    into the sin 2 theta coefficient, input.cxx:581-585), after 400 steps:

      * swimming speed 2 B1 / 3 along the axis, whatever B2 (Lighthill 1952,
        Blake 1971);
      * spin -C1 / a about the axis from the azimuthal slip (reciprocal
        theorem: Omega = -(3 / 8 pi a^3) int n x u_s dS);
      * the difference of the two flows is Blake's B2 mode with B2 = 2 B1:
        u_r = (B2 / 2)(a^4/r^4 - a^2/r^2)(3 cos^2 theta - 1), sampled at
        r = 8 = L/4 on the axis and the equator (4 points each side).

    Returns the ratios measured / predicted."""
    pp, up = squirmer_run(1.0)
    pm, um = squirmer_run(-1.0)
    U0, W0 = 2.0 * B1 / 3.0, -C1 / A
    speed = [pp["v"][2] / U0, pm["v"][2] / U0]
    spin = [pp["omega"][2] / W0, pm["omega"][2] / W0]
    du, B2 = up - um, 2.0 * B1
    stress = []
    for off in [(0, 0, 8), (0, 0, -8), (8, 0, 0), (-8, 0, 0), (0, 8, 0), (0, -8, 0)]:
        g = np.array([16, 16, 16]) + np.array(off)
        rv = g - pp["x"]
        rr = np.linalg.norm(rv)
        c = rv[2] / rr
        pred = 0.5 * B2 * (A ** 4 / rr ** 4 - A ** 2 / rr ** 2) * (3 * c * c - 1)
        stress.append(du[g[0], g[1], g[2]] @ (rv / rr) / pred)
    side = max(np.abs(pp["v"][:2]).max(), np.abs(pm["v"][:2]).max()) / U0
    tilt = max(np.abs(pp["omega"][:2]).max(), np.abs(pm["omega"][:2]).max()) / abs(W0)
    return speed, spin, stress, side, tilt


def squirmer_physics_ok(res):
    """Bands: the periodic images of a force-free swimmer (and the flow not
    yet steady at 400 steps) move the speed by a few %, the SPM interface of
    width xi moves speed and spin by up to 15 % (its effective radius is not
    a); the images of the stresslet add up to +40 % to Blake's unbounded
    field at r = L/4."""
    speed, spin, stress, side, tilt = res
    return (all(0.85 < s < 1.15 for s in speed) and all(0.75 < s < 1.05 for s in spin)
            and all(0.8 < s < 1.6 for s in stress) and side < 1e-9 and tilt < 1e-9)


def test_squirmer_speed_spin_and_stresslet():
    res = check_squirmer_physics()
    assert squirmer_physics_ok(res), res


# two squirmers on different axes (one with a rotlet) and a passive sphere,
# interfaces apart: Calc_f_slip_correct_precision (md_force.cxx:325-382)
# weights the slip field by phi_i, not phi_i / max(phi_sum, 1), so where
# interfaces overlap the particles take sum_i phi_i of the slip field and
# the fluid loses only min(phi_sum, 1) of it, and the balance holds only
# up to that overlap term
XQ = np.array([[8.5, 16.2, 15.6], [19.3, 16.7, 16.1], [14.1, 27.6, 16.4]])


def check_squirmer_momentum(steps=4):
    """The per-step balance of the plain step's momentum pin holds through the
    slip loop (sp_3d_ns.cxx:121-161) too: the slip field carries no net
    momentum (the droplet counter-flow, operate_surface.cxx:155-198), the
    particle takes the slip force of the projected field the fluid takes,
    and u leaves the loop as u* + that field."""
    from kapsel3_amd._abi import JAX_X
    sps = [make_species(mass_ratio=RHO_P, propulsion=JP_SLIP, janus_axis=JAX_Z, slip_vel=0.02, slip_mode=1.5),
           make_species(mass_ratio=RHO_P, propulsion=JP_SLIP, janus_axis=JAX_X, slip_vel=0.015, slip_mode=-0.8,
                        rotlet_c1=0.01),
           make_species(mass_ratio=RHO_P)]
    P = Params(N, a=A, a_xi=XI, dt=0.05, rotation=1, fix_cell=(0, 0, 0), lj_truncate=-1, species=sps,
               slip_iter_max=4, slip_tol=1e-3)
    r = kref.Run(P, 3)
    p = new_particles(3)
    p["x"] = p["x_nopbc"] = p["x_previous"] = XQ
    p["spec"] = [0, 1, 2]
    p["v"] = VM
    p["omega"] = OMM
    r.set_particles(p)
    r.init_fluid()
    worst = 0.0
    for _ in range(steps):
        a = r.get_particles()
        _, _, uka = r.get_fluid()
        r.step(1)
        b = r.get_particles()
        _, _, ukb = r.get_fluid()
        W, R = weights(b["x"])
        dv, dom = b["v"] - a["v"], b["omega"] - a["omega"]
        lhs = MASS * dv.sum(0) + (ukb - uka)
        rhs = (W[:, None] * dv).sum(0) + np.cross(dom, R).sum(0)
        worst = max(worst, np.abs(lhs - rhs).max() / (MASS * np.abs(dv).max()))
    return worst


def test_momentum_balance_through_the_slip_loop():
    assert check_squirmer_momentum() < 1e-11


# ------------------------------------------------------------- planted mistakes
PLAIN_CHECKS = {
    "one_step": lambda: check_one_step()[0] < 1e-12,
    "two_steps": lambda: check_two_steps((0, 0, 0))[0] < 1e-10,
    "two_steps_fix": lambda: check_two_steps((1, 0, 1))[0] < 1e-10,
    "momentum": lambda: check_momentum_balance(3) < 1e-11,
    "galilean": lambda: check_galilean(3)[0] < 1e-13,
}


@pytest.mark.parametrize("name", ["fh_pre_ns_u", "omega_r_sign", "fh_no_max", "up_no_max", "no_fix_cell",
                                  "ab2_fh_weight", "no_euler_start", "fh_old_x", "fh_sign", "ukdc_mean"])
def test_planted_step_mistake_fails_a_composite_pin(name):
    assert all(c() for c in PLAIN_CHECKS.values())  # every pin passes without it
    with kref.mutation(name):
        failed = [k for k, c in PLAIN_CHECKS.items() if not c()]
    assert failed, f"planted mistake {name} passes every composite pin"


def test_planted_advection_sign_fails_the_advection_pin():
    """The advection term is nonlinear and drops out of the composite pins
    above (they run where it is zero or below round-off); a sign slip in one
    of its products fails the random-flow brute force of the NS step."""
    import test_oracle_pins as pins
    with kref.mutation("adv_sign"):
        with pytest.raises(AssertionError):
            pins.test_ns_solver_random_flow_against_spectral_brute_force()


@pytest.mark.parametrize("name,check", [("slip_force_sign", "momentum"), ("slip_final_ustar", "momentum"),
                                        ("slip_no_reset", "momentum"), ("slip_rotlet", "physics"),
                                        ("slip_b2_full", "physics")])
def test_planted_slip_mistake_fails_a_squirmer_pin(name, check):
    with kref.mutation(name):
        if check == "momentum":
            assert not check_squirmer_momentum(2) < 1e-11, name
        else:
            assert not squirmer_physics_ok(check_squirmer_physics()), name
