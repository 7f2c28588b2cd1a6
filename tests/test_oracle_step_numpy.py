"""The oracle's whole step against an independent numpy restatement of
Time_evolution_hydro (sp_3d_ns.cxx:41-194), written from the reference
sources with numpy's FFT, so the composite order (which state each function
sees: u* after the slaved Euler step, F_H at the new position with the old
velocity, u_p with the new one, f = u_p - phi u*, zeta from u* + P f) is
pinned by a second implementation and not only through the functions it is
built from. Two spheres with overlapping interfaces in 32^3, gravity,
rotation, FIX_CELL, a random band-limited flow; one Euler step and two AB2
steps. Each term of the composition is checked to matter (a changed term
fails the comparison)."""
import math

import numpy as np

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import Params, make_species

N, A, XI, DX, RHO, ETA, DT, G = 32, 4.0, 2.0, 1.0, 1.0, 0.8, 0.05, 1e-3
MASS_RATIO = (1.5, 2.0)


def kgrid():
    i = np.arange(N)
    kint = np.where(i > N // 2, i - N, i)                       # Calc_KX/KY (fft_wrapper.cxx:45-76)
    w = 2 * np.pi / (N * DX)
    KX = (kint * w)[:, None, None] * np.ones((1, N, N // 2 + 1))
    KY = (kint * w)[None, :, None] * np.ones((N, 1, N // 2 + 1))
    KZ = (np.arange(N // 2 + 1) * w)[None, None, :] * np.ones((N, N, 1))
    KXi = kint[:, None, None] * np.ones((1, N, N // 2 + 1), dtype=int)
    KYi = kint[None, :, None] * np.ones((N, 1, N // 2 + 1), dtype=int)
    trn = (N + 2) // 3                                           # input.cxx:448
    keep = (np.abs(KXi) < trn) & (np.abs(KYi) < trn) & (np.arange(N // 2 + 1)[None, None, :] < trn)
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    IK2 = np.where(K2 > 0, 1.0 / np.where(K2 > 0, K2, 1.0), 0.0)
    return KX, KY, KZ, KXi, KYi, keep, K2, IK2


KX, KY, KZ, KXI, KYI, KEEP, K2, IK2 = kgrid()
CASE_X, CASE_Y = KXI != 0, (KXI == 0) & (KYI != 0)


def zeta2uhat(z0, z1, ukdc):
    """Zeta_k2u_k (fft_wrapper.cxx:591-688): omega from zeta by the case
    rule and div omega = 0, then u = i k x omega / k^2; u(0) = uk_dc."""
    with np.errstate(divide="ignore", invalid="ignore"):
        o0 = np.where(CASE_X, -(KY * z0 + KZ * z1) / np.where(CASE_X, KX, 1), np.where(CASE_Y, z1, z0))
        o1 = np.where(CASE_X, z0, np.where(CASE_Y, -(KZ / np.where(CASE_Y, KY, 1)) * z0, z1))
        o2 = np.where(CASE_X, z1, np.where(CASE_Y, z0, 0.0))
    u = [1j * (KY * o2 - KZ * o1) * IK2, 1j * (KZ * o0 - KX * o2) * IK2, 1j * (KX * o1 - KY * o0) * IK2]
    for d in range(3):
        u[d][0, 0, 0] = ukdc[d]
    return u


def reduce_case(w0, w1, w2):
    """Omega_k2zeta_k (fft_wrapper.cxx:227-263)."""
    return np.where(CASE_X, w1, np.where(CASE_Y, w2, w0)), np.where(CASE_X, w2, np.where(CASE_Y, w0, w1))


def fwd(a):
    return np.fft.rfftn(a)


def inv(a):
    return np.fft.irfftn(a, s=(N, N, N), axes=(0, 1, 2))


def sekibun():
    nh = int(A) + int(XI) + 2
    r = np.arange(-nh, nh + 1)
    c = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    return c[np.sqrt(((c * DX) ** 2).sum(1)) < A * DX + 0.5 * XI * DX + 1.7321 * DX]   # profile.cxx:22-65


CELLS = sekibun()


def phi_of(x):
    """Phi (profile.h:30-50)."""
    def H(y):
        return np.where(y > 0, np.exp(-(DX / np.where(y > 0, y, 1.0)) ** 2), 0.0)
    a = H(A + 0.5 * XI - x)
    return a / (a + H(x - A + 0.5 * XI))


def points(xp):
    """Particle_cell + Relative_coord (make_phi.h:114-170) and the
    minimum-image Distance of each Sekibun point (interaction.h:113-132)."""
    xi = np.floor(xp / DX).astype(int)
    res = (xp / DX - xi) * DX
    rm = (xi + CELLS + N) % N
    r = CELLS * DX - res
    d = rm * DX - xp
    q = d / (N * DX)
    d = d - np.where(q >= 0, np.floor(q + 0.5), -np.floor(-q + 0.5)) * (N * DX)
    return rm, r, phi_of(np.sqrt((d ** 2).sum(1)))


def fluid_ns(z0, z1, ukdc):
    """The slaved-Euler NS step, FIX_CELL and u* (sp_3d_ns.cxx:41-57)."""
    nu = ETA / RHO
    # NS_solver_slavedEuler (fluid_solver.cxx:348-374) via Zeta_k2advection_k
    u = [inv(c) for c in zeta2uhat(z0 * KEEP, z1 * KEEP, ukdc)]
    p = [fwd(u[1] * u[2]), fwd(u[2] * u[0]), fwd(u[0] * u[1]),
         fwd((u[1] - u[2]) * (u[1] + u[2])), fwd((u[2] - u[0]) * (u[2] + u[0]))]
    k1, k2, k3 = KX, KY, KZ
    u12 = -p[4] - p[3]
    w = [-(k2 * k3 * p[3] + (k3 - k2) * (k3 + k2) * p[0] + k3 * k1 * p[2] - k1 * k2 * p[1]),
         -(k3 * k1 * p[4] + (k1 - k3) * (k1 + k3) * p[1] + k1 * k2 * p[0] - k2 * k3 * p[2]),
         -(k1 * k2 * u12 + (k2 - k1) * (k2 + k1) * p[2] + k2 * k3 * p[1] - k3 * k1 * p[0])]  # operate_omega.cxx:57-82
    f0, f1 = reduce_case(*w)
    e = np.exp(-nu * DT * K2) - 1.0
    z0 = np.where(KEEP, z0 + e * (z0 - (1 / nu) * IK2 * f0), z0)
    z1 = np.where(KEEP, z1 + e * (z1 - (1 / nu) * IK2 * f1), z1)
    ukdc = np.zeros(3)                                            # FIX_CELL x, y, z (sp_3d_ns.cxx:46-52)
    z0, z1 = z0 * KEEP, z1 * KEEP
    us = [inv(c) for c in zeta2uhat(z0, z1, ukdc)]                # u*
    return z0, z1, ukdc, us


def project(f):
    """Solenoidal_u (operate_omega.h:105-112): FFT, 2/3 rule, k.f = 0."""
    fh = [fwd(f[d]) * KEEP for d in range(3)]
    kf = (KX * fh[0] + KY * fh[1] + KZ * fh[2]) * IK2
    return [inv(fh[0] - KX * kf), inv(fh[1] - KY * kf), inv(fh[2] - KZ * kf)]


def u2zeta(u):
    """U2zeta_k (operate_omega.h:208-211, fft_wrapper.cxx:407-476)."""
    uh = [fwd(u[d]) for d in range(3)]
    w = [1j * (KY * uh[2] - KZ * uh[1]), 1j * (KZ * uh[0] - KX * uh[2]), 1j * (KX * uh[1] - KY * uh[0])]
    z0, z1 = reduce_case(*w)
    return z0, z1, np.array([uh[d][0, 0, 0].real for d in range(3)])


def scatter_phi(P):
    """Make_phi_particle_sum (make_phi.cxx:214-260): phi_sum and the points."""
    phis = np.zeros((N, N, N))
    pts = []
    for q in P:
        rm, r, ph = points(q["x"])
        np.add.at(phis, (rm[:, 0], rm[:, 1], rm[:, 2]), ph)
        pts.append((rm, r, ph))
    return phis, pts


def step(z0, z1, ukdc, P, ts):
    z0, z1, ukdc, us = fluid_ns(z0, z1, ukdc)
    # positions (particle_solver.cxx:39-85): Euler first, then AB2
    for q in P:
        dx = DT * q["v"] if ts == 0 else 0.5 * DT * (3 * q["v"] - q["v_old"])
        q["x_nopbc"] = q["x_nopbc"] + dx
        q["x"] = np.fmod(q["x"] + dx + N * DX, N * DX)
    phis, pts = scatter_phi(P)
    phi = np.minimum(phis, 1.0)
    # F_H with u* at the new position and the old velocity (md_force.cxx:383-506)
    coef = -DX ** 3 * RHO / DT
    for q, (rm, r, ph) in zip(P, pts):
        w_ = ph / np.maximum(phis[rm[:, 0], rm[:, 1], rm[:, 2]], 1.0)
        vr = q["v"] + np.cross(q["omega"], r)
        uu = np.stack([us[d][rm[:, 0], rm[:, 1], rm[:, 2]] for d in range(3)], 1)
        fp = (vr - uu) * w_[:, None]
        q["fh"] = coef * fp.sum(0)
        q["th"] = coef * np.cross(r, fp).sum(0)
    # Force (gravity only) and the velocity update (particle_solver.cxx:103-142, 239-281)
    for q, mr in zip(P, MASS_RATIO):
        q["fr"] = q["fr"].copy()
        q["fr"][2] -= G * RHO * 4 / 3 * math.pi * A ** 3 * (mr - 1.0)   # md_force.cxx:306-324
        mass = mr * RHO * 4 / 3 * math.pi * A ** 3
        moi = 0.4 * mass * A ** 2
        q["v_old"], q["omega_old"] = q["v"].copy(), q["omega"].copy()
        if ts == 0:
            q["v"] = q["v"] + DT / mass * (q["fh"] + q["fr"])
            q["omega"] = q["omega"] + DT / moi * q["th"]
        else:
            q["v"] = q["v"] + 0.5 * DT / mass * (2 * q["fh"] + q["fr"] + q["fr_prev"])
            q["omega"] = q["omega"] + 0.5 * DT / moi * (2 * q["th"])
        q["fr_prev"], q["fr"] = q["fr"], np.zeros(3)
        q["fh_prev"], q["th_prev"] = q["fh"], q["th"]
    # u_p with the new velocity (make_phi.cxx:313-354)
    up = [np.zeros((N, N, N)) for _ in range(3)]
    for q, (rm, r, ph) in zip(P, pts):
        w_ = ph / np.maximum(phis[rm[:, 0], rm[:, 1], rm[:, 2]], 1.0)
        vr = (q["v"] + np.cross(q["omega"], r)) * w_[:, None]
        for d in range(3):
            np.add.at(up[d], (rm[:, 0], rm[:, 1], rm[:, 2]), vr[:, d])
    # f = P(u_p - phi u*), u = u* + f, zeta = curl u (f_particle.h:35-84,
    # operate_omega.cxx:171-195, fft_wrapper.cxx:407-476)
    f = project([up[d] - phi * us[d] for d in range(3)])
    return u2zeta([us[d] + f[d] for d in range(3)])


def spec(a):
    return a[:, :, 0::2] + 1j * a[:, :, 1::2]


def test_whole_step_against_numpy_restatement():
    sp = [make_species(mass_ratio=m) for m in MASS_RATIO]
    P = Params(N, dx=DX, rho=RHO, eta=ETA, dt=DT, a=A, a_xi=XI, g=G, g_dir=2, rotation=1, fix_cell=(1, 1, 1),
               lj_truncate=-1, species=sp)
    # centres 6.9 apart: the interfaces overlap (phi_sum > 1 there), so the
    # phi / max(phi_sum, 1) weights of the gather and the scatter matter
    parts = workloads.particles_at(np.array([[7.3, 9.1, 11.6], [14.1, 10.3, 12.2]]), spec=np.array([0, 1]))
    parts["v"] = [[2e-3, -1e-3, 3e-3], [-1e-3, 2e-3, -2e-3]]
    parts["omega"] = [[1e-4, 2e-4, -1e-4], [-2e-4, 1e-4, 3e-4]]
    r = kref.Run(P, 2)
    r.set_particles(parts)
    rng = np.random.default_rng(3)
    for d in range(3):
        f = np.zeros(r.shape)
        f[:, :, :N] = 1e-2 * rng.standard_normal((N, N, N))
        r.set_field(f"u{d}", f)
    r.apply("u2zeta_k")                                           # a random band-limited start
    z = r.get_fluid()
    z0, z1, ukdc = spec(z[0]).copy(), spec(z[1]).copy(), z[2].copy()
    p0 = r.get_particles()
    state = [{"x": p0["x"][k].copy(), "x_nopbc": p0["x_nopbc"][k].copy(), "v": p0["v"][k].copy(),
              "v_old": p0["v_old"][k].copy(), "omega": p0["omega"][k].copy(), "fr": p0["fr"][k].copy(),
              "fr_prev": p0["fr_previous"][k].copy()} for k in range(2)]
    for ts in range(3):
        z0, z1, ukdc = step(z0, z1, ukdc, state, ts)
    r.step(3)
    zo = r.get_fluid()
    po = r.get_particles()
    for a, b in ((spec(zo[0]), z0), (spec(zo[1]), z1)):
        s = np.abs(b[KEEP]).max()
        assert np.abs(a[KEEP] - b[KEEP]).max() < 1e-11 * s
    assert np.abs(zo[2] - ukdc).max() < 1e-11 * max(np.abs(ukdc).max(), 1e-300)
    vs = max(np.abs(q["v"]).max() for q in state)
    for k, q in enumerate(state):
        assert np.abs(po["x_nopbc"][k] - q["x_nopbc"]).max() < 1e-12 * N
        assert np.abs(po["v"][k] - q["v"]).max() < 1e-10 * vs
        assert np.abs(po["omega"][k] - q["omega"]).max() < 1e-10 * vs / A
        fs = np.abs(q["fh_prev"]).max()
        assert np.abs(po["f_hydro_previous"][k] - q["fh_prev"]).max() < 1e-10 * fs
        assert np.abs(po["torque_hydro_previous"][k] - q["th_prev"]).max() < 1e-10 * fs * A
        assert np.abs(po["fr_previous"][k] - q["fr_prev"]).max() < 1e-12 * np.abs(q["fr_prev"]).max()
    r.close()
