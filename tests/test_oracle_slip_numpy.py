"""The oracle's squirmer step against an independent numpy restatement of
Time_evolution_hydro with SW_JANUS_SLIP (sp_3d_ns.cxx:41-194, the slip loop
121-161), written from the reference sources: the mass and shell moments
(operate_surface.cxx:12-153), the droplet counter-flow solve (155-198), the
slip force field (200-363), its projection, the slip force gather
(md_force.cxx:325-382), the iterated velocity update with relaxation 0.7 and
the reset between iterations (particle_solver.cxx:144-237,
operate_surface.h:72-121), and the orientation update that the squirmer
frame reads (particle_rotation_solver.h:23-81). This pins the slip loop's
composite order: which velocity each iteration's slip profile uses, what is
reset, and which field the final constraint starts from."""
import math

import numpy as np

from oracle import kref
from kapsel3_amd import workloads
from kapsel3_amd._abi import JAX_Y, JAX_Z, JP_NONE, JP_SLIP, Params, make_species

from test_oracle_step_numpy import (A, DT, DX, ETA, KEEP, N, RHO, XI, fluid_ns, points, project, scatter_phi,
                                    spec, u2zeta)

EPS = np.finfo(float).eps
# species: (propulsion, janus axis, B1, B2/B1 as given, C1, C2, mass ratio)
SPECIES = [(JP_SLIP, JAX_Z, 0.01, 1.3, 2e-3, -1e-3, 1.0), (JP_SLIP, JAX_Y, 0.01, -0.8, 0.0, 0.0, 1.0),
           (JP_NONE, None, 0.0, 0.0, 0.0, 0.0, 1.5)]
TOL, MAXIT = 0.005, 4


def qmul(a, b):
    """qtn_prod (quaternion.h:160-168), (s, v) layout."""
    s = a[0] * b[0] - a[1:] @ b[1:]
    return np.concatenate([[s], a[0] * b[1:] + b[0] * a[1:] + np.cross(a[1:], b[1:])])


def rotate(v, q, body2space):
    """rigid_body_rotation (rigid_body.cxx:10-24): q v q* or q* v q, for rows of v."""
    qq = q if body2space else np.array([q[0], -q[1], -q[2], -q[3]])
    s, w = qq[0], qq[1:]
    return (s * s - w @ w) * v + 2.0 * ((v @ w)[:, None] * w + s * np.cross(w, v))


def less_than_mp(a, b):
    return (b - a) > (np.maximum(np.abs(a), np.abs(b)) + 10.0) * EPS     # macro.h:164-167


def squirmer_frame(r, q, axis):
    """Squirmer_coord (make_phi.h:282-376): unit r, theta, phi vectors and
    the polar angle about the janus axis, for rows of r."""
    e = np.eye(3)
    e3 = {JAX_Z: e[2], JAX_Y: e[1]}[axis]
    rb = rotate(r, q, False)
    rb = rb / np.linalg.norm(rb, axis=1)[:, None]
    ph = np.cross(e3, rb)
    dot = rb @ e3
    pn = np.linalg.norm(ph, axis=1)
    ok = (np.abs(pn) > 10 * EPS) & less_than_mp(np.abs(dot), 1.0)
    ph = np.where(ok[:, None], ph / np.where(ok, pn, 1.0)[:, None], 0.0)
    th = np.cross(ph, rb)
    tn = np.linalg.norm(th, axis=1)
    th = np.where(ok[:, None], th / np.where(ok, tn, 1.0)[:, None], 0.0)
    theta = np.where(ok, np.arccos(np.clip(dot, -1, 1)), np.where(dot > 0, 0.0, math.pi))
    return rotate(th, q, True), rotate(ph, q, True), theta


def shell(dist):
    """DPhi_compact_sin_norm (profile.h:175-184) and the shell test |r - a| < xi/2."""
    x = A - dist
    return np.where(np.abs(x) < 0.5 * XI, np.cos(math.pi * x / XI), 0.0), less_than_mp(np.abs(dist - A), 0.5 * XI)


def point_geometry(q):
    rm, r, ph = points(q["x"])
    d = rm * DX - q["x"]
    d = d - np.where(d / (N * DX) >= 0, np.floor(d / (N * DX) + 0.5), -np.floor(-d / (N * DX) + 0.5)) * (N * DX)
    return rm, r, ph, np.sqrt((d ** 2).sum(1))


def slip_velocity(sp, r, q, theta):
    th, ph, _ = squirmer_frame(r, q["q"], sp[1])
    b1, mode, c1, c2 = sp[2], sp[3] / 2.0, sp[4], sp[5]            # janus_slip_mode is halved (input.cxx:583)
    mag = b1 * (np.sin(theta) + mode * np.sin(2 * theta))
    az = c1 * np.sin(theta) + c2 * 1.5 * np.sin(2 * theta)
    return th * mag[:, None] + ph * az[:, None]


def moments(P, us):
    """Make_particle_momentum_factor (operate_surface.cxx:12-153)."""
    for q in P:
        sp = SPECIES[q["spec"]]
        rm, r, ph, dist = point_geometry(q)
        uf = np.stack([us[d][rm[:, 0], rm[:, 1], rm[:, 2]] for d in range(3)], 1)
        q["M0"], q["M1"] = ph.sum(), (ph[:, None] * r).sum(0)
        q["M2"] = np.eye(3) * (ph * dist ** 2).sum() - np.einsum("n,ni,nj->ij", ph, r, r)
        q["SM0"], q["SM1"], q["SM2"] = 0.0, np.zeros(3), np.zeros((3, 3))
        q["dv"], q["dw"] = np.zeros(3), np.zeros(3)
        if sp[0] != JP_SLIP:
            continue
        sn, inside = shell(dist)
        m = inside
        _, _, theta = squirmer_frame(r[m], q["q"], sp[1])
        phs = (1.0 - ph[m]) * sn[m]
        usl = slip_velocity(sp, r[m], q, theta) - uf[m]
        q["SM0"] = phs.sum()
        q["dv"] = (phs[:, None] * usl).sum(0)
        q["dw"] = (phs[:, None] * np.cross(r[m], usl)).sum(0)
        q["SM1"] = (phs[:, None] * r[m]).sum(0)
        q["SM2"] = np.eye(3) * (phs * dist[m] ** 2).sum() - np.einsum("n,ni,nj->ij", phs, r[m], r[m])


def droplet(q):
    """slip_droplet (operate_surface.cxx:155-198)."""
    vp, wp, mc = q["v_slip"], q["w_slip"], q["M1"]
    im = 1.0 / q["M0"]
    dP = -(q["SM0"] * vp + np.cross(wp, q["SM1"]) + q["dv"])
    dL = -(np.cross(q["SM1"], vp) + q["SM2"] @ wp + q["dw"])
    L0 = dL - np.cross(mc, dP) * im
    LL = np.eye(3) * (-(mc @ mc) * im) + q["M2"] + im * np.outer(mc, mc)
    dw = np.linalg.inv(LL) @ L0
    return im * (dP - np.cross(dw, mc)), dw


def slip_field(P, us):
    """Make_force_u_slip_particle (operate_surface.cxx:200-363)."""
    up = [np.zeros((N, N, N)) for _ in range(3)]
    for q in P:
        sp = SPECIES[q["spec"]]
        if sp[0] != JP_SLIP:
            continue
        dv, dw = droplet(q)
        rm, r, ph, dist = point_geometry(q)
        uf = np.stack([us[d][rm[:, 0], rm[:, 1], rm[:, 2]] for d in range(3)], 1)
        f = ph[:, None] * (dv + np.cross(dw, r))
        sn, inside = shell(dist)
        m = inside
        _, _, theta = squirmer_frame(r[m], q["q"], sp[1])
        phs = (1.0 - ph[m]) * sn[m]
        f[m] += phs[:, None] * (q["v_slip"] + np.cross(q["w_slip"], r[m]) + slip_velocity(sp, r[m], q, theta)
                                - uf[m])
        for d in range(3):
            np.add.at(up[d], (rm[:, 0], rm[:, 1], rm[:, 2]), f[:, d])
    return up


def orientation(q, ts):
    """MD_solver_orientation_Euler / _AB2 (particle_rotation_solver.h:23-81)."""
    dq = 0.5 * qmul(np.concatenate([[0.0], q["omega"]]), q["q"])
    if ts == 0:
        q["q_old"] = q["q"].copy()
        qn = q["q"] + DT * dq
    else:
        dqo = 0.5 * qmul(np.concatenate([[0.0], q["omega_old"]]), q["q_old"])
        q["q_old"] = q["q"].copy()
        qn = q["q"] + 3.0 * 0.5 * DT * dq - 0.5 * DT * dqo
    q["q"] = qn / np.sqrt(qn @ qn)


def slip_step(z0, z1, ukdc, P, ts):
    z0, z1, ukdc, us = fluid_ns(z0, z1, ukdc)
    for q in P:
        dx = DT * q["v"] if ts == 0 else 0.5 * DT * (3 * q["v"] - q["v_old"])
        q["x_nopbc"] = q["x_nopbc"] + dx
        q["x"] = np.fmod(q["x"] + dx + N * DX, N * DX)
        orientation(q, ts)
    phis, pts = scatter_phi(P)
    phi = np.minimum(phis, 1.0)
    coef = -DX ** 3 * RHO / DT
    for q, (rm, r, ph) in zip(P, pts):
        w_ = ph / np.maximum(phis[rm[:, 0], rm[:, 1], rm[:, 2]], 1.0)
        uu = np.stack([us[d][rm[:, 0], rm[:, 1], rm[:, 2]] for d in range(3)], 1)
        fp = (q["v"] + np.cross(q["omega"], r) - uu) * w_[:, None]
        q["fh"], q["th"] = coef * fp.sum(0), coef * np.cross(r, fp).sum(0)
    # the slip loop (sp_3d_ns.cxx:121-161)
    moments(P, us)
    for q in P:
        q["v_slip"], q["w_slip"] = q["v"].copy(), q["omega"].copy()
    it = 0
    while True:
        up = project(slip_field(P, us))
        for q, (rm, r, ph) in zip(P, pts):                        # Calc_f_slip_correct_precision
            fu = np.stack([up[d][rm[:, 0], rm[:, 1], rm[:, 2]] for d in range(3)], 1) * ph[:, None]
            q["fs"], q["ts_"] = -coef * fu.sum(0), -coef * np.cross(r, fu).sum(0)
        u_new = [up[d] + us[d] for d in range(3)]
        for q in P:
            sp = SPECIES[q["spec"]]
            if it == 0:                                             # Force (no LJ, no gravity here)
                q["fr"] = q["fr"].copy()
            mass = sp[6] * RHO * 4 / 3 * math.pi * A ** 3
            moi = 0.4 * mass * A ** 2
            q["v_old"], q["omega_old"] = q["v"].copy(), q["omega"].copy()
            if ts == 0:
                q["v"] = q["v"] + DT / mass * (q["fh"] + q["fs"] + q["fr"])
                q["omega"] = q["omega"] + DT / moi * (q["th"] + q["ts_"])
            else:
                q["v"] = q["v"] + 0.5 * DT / mass * (2 * (q["fh"] + q["fs"]) + q["fr"] + q["fr_prev"])
                q["omega"] = q["omega"] + 0.5 * DT / moi * (2 * (q["th"] + q["ts_"]))
        it += 1
        eps = 0.0
        for q in P:
            if SPECIES[q["spec"]][0] != JP_SLIP:
                continue
            nv, nw = q["v"] @ q["v"], q["omega"] @ q["omega"]
            dv = (q["v_slip"] - q["v"]) @ (q["v_slip"] - q["v"])
            dw = (q["w_slip"] - q["omega"]) @ (q["w_slip"] - q["omega"])
            eps = max(eps, (dv / nv if nv > 10 * EPS else dv) + (dw / nw if nw > 10 * EPS else dw))
        if math.sqrt(eps) < TOL or it == MAXIT:
            for q in P:                                             # reset_Forces
                q["fr_prev"], q["fr"] = q["fr"], np.zeros(3)
                q["fh_prev"], q["th_prev"], q["fs_prev"] = q["fh"], q["th"], q["fs"]
            break
        for q in P:                                                 # relaxation, then back to v_old
            q["v_slip"] = 0.3 * q["v_slip"] + 0.7 * q["v"]
            q["w_slip"] = 0.3 * q["w_slip"] + 0.7 * q["omega"]
            q["v"], q["omega"] = q["v_old"].copy(), q["omega_old"].copy()
    # u_p with the new velocity, f = P(u_p - phi u), u += f, zeta
    upp = [np.zeros((N, N, N)) for _ in range(3)]
    for q, (rm, r, ph) in zip(P, pts):
        w_ = ph / np.maximum(phis[rm[:, 0], rm[:, 1], rm[:, 2]], 1.0)
        vr = (q["v"] + np.cross(q["omega"], r)) * w_[:, None]
        for d in range(3):
            np.add.at(upp[d], (rm[:, 0], rm[:, 1], rm[:, 2]), vr[:, d])
    f = project([upp[d] - phi * u_new[d] for d in range(3)])
    return u2zeta([u_new[d] + f[d] for d in range(3)])


def test_squirmer_step_against_numpy_restatement():
    sp = [make_species(mass_ratio=s[6], propulsion=s[0], janus_axis=s[1] if s[1] is not None else 3, slip_vel=s[2],
                       slip_mode=s[3], rotlet_c1=s[4], rotlet_c2=s[5]) for s in SPECIES]
    P = Params(N, dx=DX, rho=RHO, eta=ETA, dt=DT, a=A, a_xi=XI, g=0.0, rotation=1, fix_cell=(1, 1, 1),
               lj_truncate=-1, species=sp, slip_iter_max=MAXIT, slip_tol=TOL)
    pos = np.array([[6.3, 7.6, 8.2], [20.1, 22.4, 9.7], [12.5, 24.0, 23.3], [26.0, 9.1, 24.8]])
    parts = workloads.particles_at(pos, spec=np.array([0, 1, 2, 0]), q=workloads.random_quaternions(4, 3))
    r = kref.Run(P, 4)
    r.set_particles(parts)
    rng = np.random.default_rng(5)
    for d in range(3):
        f = np.zeros(r.shape)
        f[:, :, :N] = 5e-3 * rng.standard_normal((N, N, N))
        r.set_field(f"u{d}", f)
    r.apply("u2zeta_k")
    z = r.get_fluid()
    z0, z1, ukdc = spec(z[0]).copy(), spec(z[1]).copy(), z[2].copy()
    p0 = r.get_particles()
    state = [{"spec": int(p0["spec"][k]), "x": p0["x"][k].copy(), "x_nopbc": p0["x_nopbc"][k].copy(),
              "v": p0["v"][k].copy(), "v_old": p0["v_old"][k].copy(), "omega": p0["omega"][k].copy(),
              "omega_old": p0["omega_old"][k].copy(), "q": p0["q"][k].copy(), "q_old": p0["q_old"][k].copy(),
              "fr": p0["fr"][k].copy(), "fr_prev": p0["fr_previous"][k].copy()} for k in range(4)]
    for ts in range(2):
        z0, z1, ukdc = slip_step(z0, z1, ukdc, state, ts)
    r.step(2)
    zo, po = r.get_fluid(), r.get_particles()
    for a, b in ((spec(zo[0]), z0), (spec(zo[1]), z1)):
        assert np.abs(a[KEEP] - b[KEEP]).max() < 1e-10 * np.abs(b[KEEP]).max()
    vs = max(np.abs(q["v"]).max() for q in state)
    for k, q in enumerate(state):
        assert np.abs(po["x_nopbc"][k] - q["x_nopbc"]).max() < 1e-12 * N
        assert np.abs(po["q"][k] - q["q"]).max() < 1e-12
        assert np.abs(po["v"][k] - q["v"]).max() < 1e-9 * vs, k
        assert np.abs(po["omega"][k] - q["omega"]).max() < 1e-9 * vs / A, k
        fs = max(np.abs(q["fs_prev"]).max(), 1e-300)
        assert np.abs(po["f_slip_previous"][k] - q["fs_prev"]).max() < 1e-9 * max(fs, np.abs(q["fh_prev"]).max())
    r.close()
