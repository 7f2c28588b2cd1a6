"""Pins the oracle's step functions (oracle/kref.c) to what the mathematics
fixes: closed forms, invariants, special values, brute force and published
known-answer vectors. CPU only; each test names the reference lines the
pinned oracle function restates."""
import math
import os

import numpy as np
import pytest

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import JAX_Z, JP_MOTOR, JP_NONE, Params, make_species

HERE = os.path.dirname(os.path.abspath(__file__))


def spec(a):
    return a[:, :, 0::2] + 1j * a[:, :, 1::2]


def kvec(n, L=None):
    L = L or n
    i = np.arange(n)
    k = np.where(i > n // 2, i - n, i) * (2 * np.pi / L)
    kz = np.arange(n // 2 + 1) * (2 * np.pi / L)
    return k[:, None, None], k[None, :, None], kz[None, None, :]


# ------------------------------------------------------------------ profile
@pytest.mark.parametrize("a,count", [(4.0, 1309), (5.0, 1935)])
def test_sekibun_cell_count(a, count):
    """Particle_domain (profile.cxx:22-65): points of the (2nh+1)^3 cube
    within a + xi/2 + 1.7321 dx."""
    r = kref.Run(Params(32, a=a, a_xi=2.0), 0)
    cell = r.sekibun()
    assert len(cell) == count
    nh = int(a) + 2 + 2
    assert np.abs(cell).max() <= nh
    assert len({tuple(c) for c in cell}) == count
    assert (np.sqrt((cell.astype(float) ** 2).sum(1)) < a + 1.0 + 1.7321).all()


def test_phi_special_values():
    """Phi (profile.h:30-50): 1 inside a - xi/2, 0 beyond a + xi/2, 1/2 at a,
    phi(a+t) + phi(a-t) = 1, monotone."""
    a, xi, dx = 5.0, 2.0, 1.0
    assert kref.phi(0.0, a, dx, xi) == 1.0 and kref.phi(a - xi / 2, a, dx, xi) == 1.0
    assert kref.phi(a + xi / 2, a, dx, xi) == 0.0 and kref.phi(9.0, a, dx, xi) == 0.0
    assert kref.phi(a, a, dx, xi) == 0.5
    ts = np.linspace(0, 0.999, 50)
    for t in ts:
        assert abs(kref.phi(a + t, a, dx, xi) + kref.phi(a - t, a, dx, xi) - 1.0) < 1e-15
    vals = [kref.phi(r, a, dx, xi) for r in np.linspace(a - 1.2, a + 1.2, 200)]
    assert all(x >= y for x, y in zip(vals, vals[1:]))
    # H(x) = exp(-dx^2/x^2): at r = a + xi/2 - 0.5, phi = e^-4 / (e^-4 + e^-1/2.25)
    r = a + xi / 2 - 0.5
    h1, h2 = math.exp(-1 / 0.25), math.exp(-1 / (1.5 ** 2))
    assert abs(kref.phi(r, a, dx, xi) - h1 / (h1 + h2)) < 1e-16


# ------------------------------------------------------------------ noise
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors published with Random123
    (kat_vectors: philox4x32 10 rounds)."""
    assert list(kref.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(kref.philox([0xffffffff] * 4, [0xffffffff] * 2)) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert list(kref.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_noise_normals_statistics():
    g = np.array([kref.noise_normals(42, 0, n, 7) for n in range(20000)]).ravel()
    assert abs(g.mean()) < 0.02 and abs(g.var() - 1.0) < 0.03
    assert abs(((g - g.mean()) ** 4).mean() / g.var() ** 2 - 3.0) < 0.1  # Gaussian kurtosis
    h = np.array([kref.noise_normals(42, 1, n, 7) for n in range(2000)]).ravel()
    assert abs(np.corrcoef(g[:h.size], h)[0, 1]) < 0.05  # streams independent


def test_mt19937_reference_vectors():
    """The reference's MT19937 (mt19937ar.c) build, seed 5489: first output
    3499211612 is the generator's published check value."""
    d = np.load(os.path.join(HERE, "golden", "mt19937.npz"))
    assert int(d["int32_seed5489"][0]) == 3499211612
    assert ((d["real3_seed0"] > 0) & (d["real3_seed0"] < 1)).all()


# ------------------------------------------------------------------ spectral
def _run_fields(n=16, **kw):
    return kref.Run(Params(n, **kw), 0)


def test_shear_wave_curl_and_inverse():
    """U_k2zeta_k (fft_wrapper.cxx:407-476) and Zeta_k2u_k (591-688) on
    u = (0, 0, A sin(k1 x)): zeta0(kx=+-1) = -k1 A N^3/2, then back to u."""
    n, A = 16, 1e-3
    r = _run_fields(n)
    x = np.arange(n)
    u2 = np.zeros(r.shape)
    u2[:, :, :n] = A * np.sin(2 * np.pi * x / n)[:, None, None]
    r.set_field("u0", np.zeros(r.shape))
    r.set_field("u1", np.zeros(r.shape))
    r.set_field("u2", u2)
    r.apply("u2zeta_k")
    z0, z1, uk = r.get_fluid()
    k1 = 2 * np.pi / n
    c0 = spec(z0)
    assert abs(c0[1, 0, 0] - (-k1 * A * n ** 3 / 2)) < 1e-15
    assert abs(c0[n - 1, 0, 0] - (-k1 * A * n ** 3 / 2)) < 1e-15
    c0[1, 0, 0] = c0[n - 1, 0, 0] = 0
    assert np.abs(c0).max() < 1e-15 and np.abs(spec(z1)).max() < 1e-15
    assert np.abs(uk).max() < 1e-15
    r.apply("zeta_k2u")
    assert np.abs(r.get_field("u2")[:, :, :n] - u2[:, :, :n]).max() < 1e-17
    assert np.abs(r.get_field("u0")[:, :, :n]).max() < 1e-17


def test_two_thirds_rule_kept_fraction():
    """Truncate_two_third_rule (fft_wrapper.h:754-796) inside Solenoidal_u:
    keeps |kx|,|ky| < TRN, kz < TRN with TRN = (N+2)/3 (input.cxx:448)."""
    n = 32
    r = _run_fields(n)
    rng = np.random.default_rng(0)
    for d in range(3):
        f = np.zeros(r.shape)
        f[:, :, :n] = rng.standard_normal((n, n, n))
        r.set_field(f"up{d}", f)
    r.apply("f_particle_sole")  # phi = 0: f = P(trunc(up))
    u = r.get_field("f0")[:, :, :n]
    s = np.fft.rfftn(u)
    nz = np.abs(s) > 1e-12 * np.abs(s).max()
    trn = (n + 2) // 3
    kx, ky, kz = [np.rint(k / (2 * np.pi / n)).astype(int) for k in kvec(n)]
    kept = (np.abs(kx) < trn) & (np.abs(ky) < trn) & (kz < trn)
    assert not (nz & ~kept).any()
    # the kept set has (2 TRN - 1)^2 TRN modes
    assert kept.sum() == (2 * trn - 1) ** 2 * trn
    assert nz.sum() > 0.95 * kept.sum()


def _set_velocity(r, n, u):
    for d in range(3):
        f = np.zeros(r.shape)
        f[:, :, :n] = u[d]
        r.set_field(f"u{d}", f)
    r.apply("u2zeta_k")


def test_slaved_euler_shear_wave_decays_exactly():
    """NS_solver_slavedEuler (fluid_solver.cxx:348-374): a shear wave has
    zero advection, so zeta -> zeta exp(-nu k^2 dt)."""
    n, A, eta, dt = 16, 1e-2, 0.7, 0.05
    r = _run_fields(n, eta=eta, dt=dt)
    x = np.arange(n)[:, None, None] * np.ones((1, n, n))
    _set_velocity(r, n, [0 * x, 0 * x, A * np.sin(2 * np.pi * x / n)])
    z_before = spec(r.get_fluid()[0]).copy()
    r.apply("ns_solver")
    z_after = spec(r.get_fluid()[0])
    k1 = 2 * np.pi / n
    assert abs(z_after[1, 0, 0] / z_before[1, 0, 0] - math.exp(-eta * k1 * k1 * dt)) < 1e-14


def test_slaved_euler_taylor_green_decays_exactly():
    """Taylor-Green vortex: u.grad(u) is a gradient, so -Omega from
    U2advection_k (operate_omega.cxx:12-85) cancels and zeta decays with
    exp(-nu 2 k^2 dt)."""
    n, A, eta, dt = 16, 5e-2, 1.0, 0.05
    r = _run_fields(n, eta=eta, dt=dt)
    g = 2 * np.pi * np.arange(n) / n
    X, Y = np.meshgrid(g, g, indexing="ij")
    ux = (A * np.sin(X) * np.cos(Y))[:, :, None] * np.ones((1, 1, n))
    uy = (-A * np.cos(X) * np.sin(Y))[:, :, None] * np.ones((1, 1, n))
    _set_velocity(r, n, [ux, uy, 0 * ux])
    z1b = spec(r.get_fluid()[1]).copy()
    r.apply("ns_solver")
    z1a = spec(r.get_fluid()[1])
    k2 = 2 * (2 * np.pi / n) ** 2
    m = np.abs(z1b) > 1e-6 * np.abs(z1b).max()
    assert m.sum() >= 4
    assert np.abs(z1a[m] / z1b[m] - math.exp(-eta * k2 * dt)).max() < 1e-13


def test_ns_solver_random_flow_against_spectral_brute_force():
    """NS_solver_slavedEuler (fluid_solver.cxx:348-374) with the advection
    term of U2advection_k (operate_omega.cxx:12-85) on a random solenoidal
    flow with 2/3-truncated modes: vorticity w = ik x u evolves as
    w' = exp(-nu k^2 dt) w + (1 - exp(-nu k^2 dt)) Omega / (nu k^2) with
    Omega = -FFT(curl((u.grad) u)), (u.grad) u_i = d_j(u_i u_j) from the
    products in real space (alias-free on the kept modes). Compared through
    u' = ik x w' / k^2 plus the unchanged mean."""
    n, eta, dt = 16, 0.8, 0.05
    r = _run_fields(n, eta=eta, dt=dt)
    kx, ky, kz = kvec(n)
    K = np.stack(np.broadcast_arrays(kx, ky, kz))
    trn = (n + 2) // 3
    ki = [np.rint(k / (2 * np.pi / n)).astype(int) for k in (kx, ky, kz)]
    kept = (np.abs(ki[0]) < trn) & (np.abs(ki[1]) < trn) & (ki[2] < trn)
    rng = np.random.default_rng(11)
    k2 = (K ** 2).sum(0)
    k2s = np.where(k2 == 0, 1.0, k2)
    F = (rng.standard_normal((3,) + kept.shape) + 1j * rng.standard_normal((3,) + kept.shape)) * kept
    F = F - K * (K * F).sum(0) / k2s                                   # solenoidal
    u = np.stack([np.fft.irfftn(F[d], s=(n, n, n), axes=(0, 1, 2)) for d in range(3)])
    u = 0.3 * u / np.abs(u).max() + np.array([0.02, -0.01, 0.03])[:, None, None, None]
    _set_velocity(r, n, list(u))
    r.apply("ns_solver")
    r.apply("zeta_k2u")
    got = np.stack([r.get_field(f"u{d}")[:, :, :n] for d in range(3)])
    U = np.stack([np.fft.rfftn(u[d]) for d in range(3)])
    P = {(i, j): np.fft.rfftn(u[i] * u[j]) for i in range(3) for j in range(3)}
    adv = np.stack([sum(1j * K[j] * P[(i, j)] for j in range(3)) for i in range(3)])
    omega = -1j * np.cross(K, adv, axis=0)                             # -FFT(curl adv)
    w = 1j * np.cross(K, U, axis=0)
    e = np.exp(-eta * k2 * dt)
    w2 = (e * w + (1 - e) * omega / (eta * k2s)) * kept
    U2 = 1j * np.cross(K, w2, axis=0) / k2s
    U2[:, 0, 0, 0] = U[:, 0, 0, 0]
    want = np.stack([np.fft.irfftn(U2[d], s=(n, n, n), axes=(0, 1, 2)) for d in range(3)])
    change = np.abs(want - u).max()
    assert change > 1e-3                                                # the step does something
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    # the nonlinear term matters here: without it the answer is off
    UL = 1j * np.cross(K, e * w * kept, axis=0) / k2s
    UL[:, 0, 0, 0] = U[:, 0, 0, 0]
    lin = np.stack([np.fft.irfftn(UL[d], s=(n, n, n), axes=(0, 1, 2)) for d in range(3)])
    assert np.abs(lin - want).max() > 1e-4 * np.abs(want).max()


def test_projection_solenoidal_and_idempotent():
    """Solenoidal_u (operate_omega.h:105-112): k.f = 0, only kept modes,
    P(P f) = P f."""
    n = 16
    r = _run_fields(n)
    rng = np.random.default_rng(1)
    for d in range(3):
        f = np.zeros(r.shape)
        f[:, :, :n] = rng.standard_normal((n, n, n))
        r.set_field(f"up{d}", f)
    r.apply("f_particle_sole")  # phi = 0: f = P(up)
    f = [r.get_field(f"f{d}")[:, :, :n] for d in range(3)]
    F = [np.fft.rfftn(c) for c in f]
    kx, ky, kz = kvec(n)
    div = kx * F[0] + ky * F[1] + kz * F[2]
    assert np.abs(div).max() < 1e-12 * max(np.abs(F[0]).max(), 1)
    for d in range(3):
        full = np.zeros(r.shape)
        full[:, :, :n] = f[d]
        r.set_field(f"up{d}", full)
    r.apply("f_particle_sole")
    for d in range(3):
        assert np.abs(r.get_field(f"f{d}")[:, :, :n] - f[d]).max() < 1e-14


# ------------------------------------------------------------------ particles
def _two_species_params(n, **kw):
    return Params(n, a=4.0, a_xi=2.0, species=[make_species(mass_ratio=1.0), make_species(mass_ratio=2.0)], **kw)


def _min_image(d, L):
    return d - L * np.where(d / L >= 0, np.floor(d / L + 0.5), np.ceil(d / L - 0.5))


def test_scatter_and_gather_brute_force():
    """Make_phi_particle_sum (make_phi.cxx:214-260), Make_u_particle_sum
    (313-354) and Calc_f_hydro_correct_precision (md_force.cxx:383-506)
    against an all-points x all-particles sum, with overlapping particles
    (phi_sum > 1) and a particle wrapping through the periodic boundary."""
    n, a, xi = 24, 4.0, 2.0
    P = _two_species_params(n, dt=0.05, rotation=1)
    pos = np.array([[0.7, 23.4, 11.2], [8.1, 22.9, 12.0], [15.5, 9.25, 3.5]])
    parts = workloads.particles_at(pos, spec=np.array([0, 1, 0]))
    rng = np.random.default_rng(2)
    parts["v"] = rng.normal(0, 1e-2, (3, 3))
    parts["omega"] = rng.normal(0, 1e-2, (3, 3))
    r = kref.Run(P, 3)
    r.set_particles(parts)
    u = [rng.normal(0, 1e-2, r.shape) for _ in range(3)]
    for d in range(3):
        r.set_field(f"u{d}", u[d])
    r.apply("make_phi")
    ps = r.get_field("phi_sum")[:, :, :n]
    g = np.arange(n, dtype=float)
    G = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1)
    phis, rs = [], []
    for k in range(3):
        rv = _min_image(G - pos[k], n)
        dist = np.sqrt((rv ** 2).sum(-1))
        phis.append(np.vectorize(lambda t: kref.phi(t, a, 1.0, xi))(dist))
        rs.append(rv)
    bf = sum(phis)
    assert np.abs(ps - bf).max() < 1e-14
    assert bf.max() > 1.0
    r.apply("make_up")
    den = np.maximum(bf, 1.0)
    for d in range(3):
        bf_up = sum(phis[k] / den * (parts["v"][k, d] + np.cross(parts["omega"][k], rs[k])[..., d]) for k in range(3))
        assert np.abs(r.get_field(f"up{d}")[:, :, :n] - bf_up).max() < 1e-15
    r.apply("f_hydro")
    got = r.get_particles()
    coef = -1.0 / 0.05
    for k in range(3):
        w = (phis[k] / den).astype(np.longdouble)
        vrel = parts["v"][k] + np.cross(parts["omega"][k], rs[k]) - np.stack([u[d][:, :, :n] for d in range(3)], -1)
        fp = w[..., None] * vrel
        F = coef * fp.reshape(-1, 3).sum(0)
        T = coef * np.cross(rs[k], fp).reshape(-1, 3).sum(0)
        assert np.abs(got["f_hydro"][k] - F).max() < 1e-13 * max(np.abs(F).max(), 1e-30) + 1e-18
        assert np.abs(got["torque_hydro"][k] - T).max() < 1e-13 * max(np.abs(T).max(), 1e-30) + 1e-18


def test_integrators_closed_form():
    """MD_solver_velocity_Euler/AB2_hydro (particle_solver.cxx:103-142,
    239-281), position Euler/AB2 (39-85), orientation (particle_rotation_solver.h)
    and gravity (md_force.cxx:306-324) on one free particle."""
    n, dt, G, a = 32, 0.05, 1e-3, 4.0
    P = _two_species_params(n, dt=dt, g=G, g_dir=2, rotation=1, lj_truncate=1)
    parts = workloads.particles_at(np.array([[31.9, 5.0, 5.0]]), spec=np.array([1]))
    w = 0.3
    parts["v"] = [[4.0, 0.0, 0.0]]
    parts["omega"] = [[0.0, 0.0, w]]
    parts["f_hydro"] = [[1e-2, -2e-2, 3e-2]]
    r = kref.Run(P, 1)
    r.set_particles(parts)
    M = 2.0 * 4 / 3 * np.pi * a ** 3
    Fg = np.array([0, 0, -G * 4 / 3 * np.pi * a ** 3 * (2.0 - 1.0)])
    r.apply("position_euler")
    p = r.get_particles()
    assert abs(p["x"][0, 0] - ((31.9 + dt * 4.0) % 32)) < 1e-13 and abs(p["x_nopbc"][0, 0] - (31.9 + 0.2)) < 1e-13
    h = w * dt / 2
    assert np.abs(p["q"][0] - np.array([1, 0, 0, h]) / math.sqrt(1 + h * h)).max() < 1e-15
    r.apply("velocity_euler")
    p = r.get_particles()
    v1 = np.array([4.0, 0, 0]) + dt / M * (np.array([1e-2, -2e-2, 3e-2]) + Fg)
    assert np.abs(p["v"][0] - v1).max() < 1e-15
    assert np.abs(p["fr_previous"][0] - Fg).max() < 1e-18 and np.abs(p["f_hydro"][0]).max() == 0
    q1 = p["q"][0].copy()
    r.apply("position_ab2")
    p = r.get_particles()
    x2 = (p["x_previous"][0] + dt / 2 * (3 * v1 - np.array([4.0, 0, 0])))
    assert np.abs(p["x"][0] - np.mod(x2, 32)).max() < 1e-13

    def qdot(q, om):  # 1/2 (0, w) o q for w along z
        return 0.5 * om * np.array([-q[3], -q[2], q[1], q[0]])
    om1 = p["omega"][0][2]
    q2 = q1 + 3 * dt / 2 * qdot(q1, om1) - dt / 2 * qdot(np.array([1.0, 0, 0, 0]), w)
    assert np.abs(p["q"][0] - q2 / np.linalg.norm(q2)).max() < 1e-15
    r.apply("velocity_ab2")  # f_hydro = 0 now, fr = Fg, fr_prev = Fg
    p = r.get_particles()
    v2 = v1 + dt / (2 * M) * (Fg + Fg)
    assert np.abs(p["v"][0] - v2).max() < 1e-15


@pytest.mark.parametrize("n,mult", [(64, 1), (32, 2)])
def test_lennard_jones_pair_and_link_cell_multiplicity(n, mult):
    """Lennard_Jones_f (interaction.h:160-199) for a pair inside the cutoff
    2^(1/6) sigma; with N/16 = 2 link cells the reference visits a pair in
    different cells twice per axis (md_force.cxx:86-101)."""
    dt, a = 0.05, 4.0
    P = _two_species_params(n, dt=dt, lj_truncate=1, lj_powers=0, epsilon=1.0, lj_cell_compat=1)
    x0 = np.array([14.0, 3.0, 3.0])
    rr = 8.4  # sigma = 8, cutoff 8.98
    parts = workloads.particles_at(np.array([x0, x0 + [rr, 0, 0]]), spec=np.array([0, 0]))
    r = kref.Run(P, 2)
    r.set_particles(parts)
    r.apply("velocity_euler")
    p = r.get_particles()
    s = 8.0 / rr
    fmag = 24.0 / rr ** 2 * (2 * s ** 12 - s ** 6) * rr
    M = 4 / 3 * np.pi * a ** 3
    assert abs(p["v"][0, 0] - (-mult * fmag * dt / M)) < 1e-15
    assert abs(p["v"][1, 0] - (mult * fmag * dt / M)) < 1e-15
    # pairs beyond the cutoff feel nothing
    parts = workloads.particles_at(np.array([x0, x0 + [9.1, 0, 0]]), spec=np.array([0, 0]))
    r.set_particles(parts)
    r.apply("velocity_euler")
    assert np.abs(r.get_particles()["v"]).max() == 0.0


def test_lennard_jones_brute_force_many():
    """Link-cell-free pair sum against O(P^2) numpy, lc >= 3 (every pair once)."""
    n = 64
    P = Params(n, a=5.0, a_xi=2.0, lj_truncate=1, epsilon=1.0, species=[make_species()], dt=0.05)
    pos = workloads.place_spheres(40, (n, n, n), 5.0, gap=0.05, seed=9)
    r = kref.Run(P, len(pos))
    r.set_particles(workloads.particles_at(pos))
    r.apply("velocity_euler")
    v = r.get_particles()["v"]
    M = 4 / 3 * np.pi * 125
    F = np.zeros((len(pos), 3))
    cut = 2 ** (1 / 6) * 10
    for i in range(len(pos)):
        for j in range(len(pos)):
            if i == j:
                continue
            d = _min_image(pos[i] - pos[j], n)
            rij = np.sqrt((d ** 2).sum())
            if rij < cut:
                s = 10 / rij
                F[i] += 24 / rij ** 2 * (2 * s ** 12 - s ** 6) * d
    assert np.abs(F).max() > 0
    assert np.abs(v - 0.05 / M * F).max() < 1e-13 * np.abs(0.05 / M * F).max()


def test_single_sphere_settles_at_periodic_array_drag():
    """Whole step, one driven sphere at 32^3: the settling speed approaches Hasimoto's
    periodic-array value (U/U_St = 1 - 1.7601 c^1/3 + c - 1.5593 c^2)."""
    P, parts = workloads.plain_single(32)
    r = kref.Run(P, 1)
    r.set_particles(parts)
    r.init_fluid()
    r.step(1001)
    v = r.get_particles()["v"][0]
    U = workloads.hasimoto_velocity(1e-3, 4.0, 1.0, 32 ** 3)
    assert 0.90 < v[2] / U < 1.02
    assert np.abs(v[:2]).max() < 1e-12 * abs(v[2])
    assert r.get_time().ts == 1001
