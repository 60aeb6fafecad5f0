"""The oracle's momentum time steps (oracle/timestep.py) and advective term, pinned on the CPU.

  * mk_advective_m_fluxdiv: second-order convergence to the analytic -div(rho u u) of smooth
    periodic fields (2-D and 3-D, variable density), discrete conservation (the face sums
    telescope to zero on a periodic box), zero update on wall faces;
  * advance_timestep_inertial: a shear mode u = (A sin 2 pi k y, 0, 0) has no advection and no
    pressure, so the step multiplies it by the Crank-Nicolson factor (rho/dt - eta mu/2) /
    (rho/dt + eta mu/2), mu = 4/h^2 sin^2(pi k h) the eigenvalue of the discrete Laplacian;
    a Taylor-Green vortex translated at U0 (an exact Navier-Stokes solution whose advective
    term is not a gradient) converges at second order in the velocity and at first order in
    the pressure -- a dropped 1/2 on the advective term fails both;
  * advance_timestep_overdamped: with forcings F1 and F2 = F a shear mode, the corrector's
    Stokes solve gives v = F / (eta mu) whatever F1 was (up to the velocity mean, which is
    the null space of the theta = 0 system on a periodic box and stays that of the input).
"""
import numpy as np
import pytest

from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, Grid

TP = 2 * np.pi


def _oracle(g, vt=1, **kn):
    from oracle import Oracle

    kn.setdefault("gmres_max_inner", 30)
    kn.setdefault("gmres_rel_tol", 1e-12)
    return Oracle(g, visc_type=vt, **kn)


def _adv_fields(g):
    U = [lambda X, Y, Z: np.sin(TP * Y) * np.cos(TP * Z) + 0.3, lambda X, Y, Z: np.sin(TP * X) * np.cos(TP * Z),
         lambda X, Y, Z: np.cos(TP * X) * np.sin(TP * Y)]
    rho = lambda X, Y, Z: 1 + 0.3 * np.sin(TP * (X + Z)) * np.cos(TP * Y)  # noqa: E731
    u, r, ex = [], [], []
    h = 1e-5
    for a in range(g.dm):
        c = g.face_coords(a)
        X, Y = c[0], c[1]
        Z = c[2] if g.dm == 3 else np.zeros_like(X)
        u.append(U[a](X, Y, Z))
        r.append(rho(X, Y, Z))
        e = np.zeros_like(X)
        for d in range(g.dm):
            s = [h if q == d else 0.0 for q in range(3)]
            fp = rho(X + s[0], Y + s[1], Z + s[2]) * U[a](X + s[0], Y + s[1], Z + s[2]) * U[d](X + s[0], Y + s[1], Z + s[2])
            fm = rho(X - s[0], Y - s[1], Z - s[2]) * U[a](X - s[0], Y - s[1], Z - s[2]) * U[d](X - s[0], Y - s[1], Z - s[2])
            e -= (fp - fm) / (2 * h)
        ex.append(e)
    return u, r, ex


@pytest.mark.parametrize("dm", [2, 3])
def test_advective_converges_second_order(dm):
    errs = []
    for n in (16, 32, 64) if dm == 2 else (12, 24, 48):
        g = Grid((n,) * dm, 1.0 / n)
        u, r, ex = _adv_fields(g)
        A = _oracle(g).advective_m_fluxdiv(u, r)
        errs.append(max(np.abs(A[a] - ex[a]).max() for a in range(dm)))
    print(dm, errs)
    assert errs[0] / errs[1] > 3.3 and errs[1] / errs[2] > 3.7  # 3-D at 12^3 is still pre-asymptotic (3.43)


def test_advective_conserves_and_zeroes_walls():
    g = Grid((16, 12, 10), 1 / 16)
    rng = np.random.default_rng(1)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(3)]
    r = [1 + 0.2 * rng.random(g.face_shape(a)) for a in range(3)]
    for a in range(3):  # periodic data: the duplicated face n holds face 0's value
        for x in (u[a], r[a]):
            np.moveaxis(x, 2 - a, 0)[-1] = np.moveaxis(x, 2 - a, 0)[0]
    A = _oracle(g).advective_m_fluxdiv(u, r)
    for a in range(3):
        # every face once (the duplicated periodic face n is the same degree of freedom as 0)
        ax = 2 - a
        s = np.take(A[a], range(g.face_shape(a)[ax] - 1), axis=ax).sum()
        assert abs(s) < 1e-11 * np.abs(A[a]).sum()
    gw = Grid((8, 8, 8), 1 / 8, bc_lo=(PERIODIC, NO_SLIP_WALL, PERIODIC), bc_hi=(PERIODIC, NO_SLIP_WALL, PERIODIC))
    u = [rng.standard_normal(gw.face_shape(a)) for a in range(3)]
    r = [np.ones(gw.face_shape(a)) for a in range(3)]
    A = _oracle(gw).advective_m_fluxdiv(u, r)
    assert np.all(A[1][:, 0, :] == 0) and np.all(A[1][:, -1, :] == 0) and np.abs(A[1][:, 1:-1, :]).max() > 0


def _uniform(g, eta):
    rho = [np.ones(g.face_shape(a)) for a in range(g.dm)]
    return (rho, eta * np.ones(g.cc_shape()), [eta * np.ones(g.edge_shape(e)) for e in range(g.nedge)],
            np.zeros(g.cc_shape()))


@pytest.mark.parametrize("dm", [2, 3])
def test_inertial_shear_mode_crank_nicolson(dm):
    from oracle import timestep as TS

    n, k, A, eta, dt = 16, 2, 0.7, 0.3, 0.01
    g = Grid((n,) * dm, 1.0 / n)
    Y = g.face_coords(0)[1]
    u = [A * np.sin(TP * k * Y)] + [np.zeros(g.face_shape(a)) for a in range(1, dm)]
    p = np.zeros(g.cc_shape())
    un, pn, s1, s2 = TS.advance_inertial(_oracle(g), g, dt, *_uniform(g, eta), u, p, gmres_rel_tol=1e-13)
    mu = 4 * n * n * np.sin(np.pi * k / n) ** 2
    fac = (1 / dt - 0.5 * eta * mu) / (1 / dt + 0.5 * eta * mu)
    assert np.abs(un[0] - fac * u[0]).max() < 1e-12 * A
    assert max(np.abs(x).max() for x in un[1:]) < 1e-12 and np.abs(pn).max() < 1e-10


def _translated_tg(g, t, nu, U0=1.0):
    F = np.exp(-2 * TP ** 2 * nu * t)
    X, Y = g.face_coords(0)[:2]
    u = U0 + np.sin(TP * (X - U0 * t)) * np.cos(TP * Y) * F
    X, Y = g.face_coords(1)[:2]
    v = -np.cos(TP * (X - U0 * t)) * np.sin(TP * Y) * F
    Xc, Yc = g.cc_coords()[:2]
    p = 0.25 * (np.cos(2 * TP * (Xc - U0 * t)) + np.cos(2 * TP * Yc)) * F ** 2
    return [u, v], p


def test_inertial_translated_taylor_green_convergence():
    from oracle import timestep as TS

    nu, T = 0.05, 0.1
    errs = []
    for n in (16, 32, 64):
        g = Grid((n, n), 1.0 / n)
        dt = 0.2 / n
        o = _oracle(g)
        coef = _uniform(g, nu)
        u, p = _translated_tg(g, 0.0, nu)
        for _ in range(int(round(T / dt))):
            u, p, _, _ = TS.advance_inertial(o, g, dt, *coef, u, p, gmres_rel_tol=1e-12)
        ue, pe = _translated_tg(g, T, nu)
        errs.append((max(np.abs(u[a] - ue[a]).max() for a in range(2)),
                     np.abs((p - p.mean()) - (pe - pe.mean())).max()))
    print(errs)
    assert errs[0][0] / errs[1][0] > 3.7 and errs[1][0] / errs[2][0] > 3.8
    assert errs[0][1] / errs[1][1] > 1.8 and errs[1][1] / errs[2][1] > 1.8 and errs[2][1] < 5e-3


@pytest.mark.parametrize("dm", [2, 3])
def test_overdamped_forced_shear_mode(dm):
    from oracle import timestep as TS

    n, k, eta = 16, 2, 0.3
    g = Grid((n,) * dm, 1.0 / n)
    Y = g.face_coords(0)[1]
    F = [np.sin(TP * k * Y)] + [np.zeros(g.face_shape(a)) for a in range(1, dm)]
    rng = np.random.default_rng(2)
    u0 = [rng.standard_normal(g.face_shape(a)) for a in range(dm)]
    un, pn, s1, s2 = TS.advance_overdamped(_oracle(g), g, *_uniform(g, eta), u0, np.zeros(g.cc_shape()),
                                           f1=[2 * x for x in F], f2=F, gmres_rel_tol=1e-13)
    mu = 4 * n * n * np.sin(np.pi * k / n) ** 2
    ref = F[0] / (eta * mu)
    # with theta = 0 on a periodic box the velocity mean is the null space: the increments have
    # zero mean, so each component keeps the mean of u0 and matches the closed form up to it
    for a, x in enumerate(un):
        d = x - (ref if a == 0 else 0.0)
        assert d.max() - d.min() < 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("wd", [0, 1])
@pytest.mark.parametrize("dm", [2, 3])
def test_steps_hold_wall_faces(dm, wd):
    """Wall-normal faces of v are zero after either step (multifab_physbc_domainvel at entry,
    homogeneous walls in both solves), in 2-D and 3-D, whatever the input held there."""
    from oracle import timestep as TS

    n = (12, 10) if dm == 2 else (8, 6, 10)
    walls = [PERIODIC] * dm
    walls[wd] = NO_SLIP_WALL
    g = Grid(n, 1 / 12, bc_lo=tuple(walls), bc_hi=tuple(walls))
    rng = np.random.default_rng(4)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(dm)]
    p = rng.standard_normal(g.cc_shape())
    for step in ("inertial", "overdamped"):
        o = _oracle(g, vt=-1, gmres_rel_tol=1e-10)
        if step == "inertial":
            un, _, _, _ = TS.advance_inertial(o, g, 0.05, *_uniform(g, 1.0), u, p, gmres_rel_tol=1e-10)
        else:
            f = [rng.standard_normal(g.face_shape(a)) for a in range(dm)]
            un, _, _, _ = TS.advance_overdamped(o, g, *_uniform(g, 1.0), u, p, f1=f, f2=f, gmres_rel_tol=1e-10)
        w = np.moveaxis(un[wd], 2 - wd, 0)
        assert np.all(w[0] == 0) and np.all(w[-1] == 0)
        assert np.abs(w[1:-1]).max() > 1e-3
