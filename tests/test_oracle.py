"""Pins of the FP64 oracle (CPU).  The oracle is the checker of the HIP path, so each of its
parts is tied to something outside itself:

* stag_applyop / compute_grad / compute_div / averaging: the reference's own truncation-error
  goldens (Accuracy_tests/{2d,3d}_dis_accuracy.dat, every stored visc_type, 2-D and 3-D);
* staggered multigrid (smoother, transfers, coarsening): V-cycle contraction against the
  reference's stored history (Mg_StagMg_performance/3d_periodic_precon_finegridGS_vel.dat);
* GMRES + preconditioner: a dense direct solve of the assembled Stokes matrix on a tiny grid;
* MAC multigrid: converges to the dense solution of D(alpha^-1 G) phi = r;
* stochastic_m_fluxdiv: single-impulse stencils, wall rules (sqrt 2 / 0), telescoping sums.
"""
import json
import os

import numpy as np
import pytest

import oracle
from fixtures import accuracy_problem
from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "accuracy_goldens.json")))


def _golden_id(r):
    return f"{r['dim']}d-n{r['n']}-vt{r['visc_type']}-pc{r['prob_coeff']}"


def _check_golden(r):
    n, dim = r["n"], r["dim"]
    g = Grid((n,) * dim, 1.0 / n)
    pr = accuracy_problem(g, r["visc_type"], r["prob_coeff"])
    o = oracle.Oracle(g, visc_type=r["visc_type"])
    o.set_coeffs(pr["alpha"], pr["beta"], pr["beta_ed"], pr["gamma"], 0.0)
    bu, bp = o.apply_matrix(pr["u"], pr["p"])
    errs = [np.abs(pr["f"][d] - bu[d]).max() for d in range(dim)] + [np.abs(pr["fp"] - bp).max()]
    # the goldens are differences of O(n^2)-sized stencil terms: an absolute rounding bound of
    # 32 eps n^2 (a wrong coefficient, sign or index changes them by O(1) relative instead)
    tol = 32 * np.finfo(float).eps * n * n
    for e, c in zip(errs, "UVW"[:dim] + "P"):
        assert abs(e - r[c]) <= tol, (c, e, r[c])


@pytest.mark.parametrize("r", [r for r in GOLD if r["n"] <= 64], ids=_golden_id)
def test_accuracy_golden(r):
    _check_golden(r)


@pytest.mark.parametrize("r", [r for r in GOLD if r["n"] == 128 and (r["dim"] == 2 or r["visc_type"] in (1, -1))],
                         ids=_golden_id)
def test_accuracy_golden_128(r):
    _check_golden(r)


def test_stag_mg_contraction_matches_reference_history():
    """Reference (3-D periodic, visc_type 1, colour GS, 2/2/8 sweeps, theta = 0): residual ratios
    per V-cycle 5.59e-11 -> 5.33e-12 -> 5.21e-13 -> 5.19e-14 at cycles 8..11, i.e. a factor of
    ~10.3 per cycle.  Same settings here on a 32^3 grid from a zero-mean random rhs."""
    n = 32
    g = Grid((n, n, n), 1.0)
    o = oracle.Oracle(g, visc_type=1, stag_minwidth=4)
    one = lambda s: np.ones(s)  # noqa: E731
    o.set_coeffs([one(g.face_shape(d)) for d in range(3)], one(g.cc_shape()),
                 [one(g.edge_shape(e)) for e in range(3)], one(g.cc_shape()), 0.0)
    rng = np.random.default_rng(10)
    rhs = [rng.random(g.face_shape(d)) for d in range(3)]
    rhs[0][..., -1] = rhs[0][..., 0]
    rhs[1][:, -1, :] = rhs[1][:, 0, :]
    rhs[2][-1] = rhs[2][0]
    uq = [rhs[0][..., :-1], rhs[1][:, :-1, :], rhs[2][:-1]]
    for d in range(3):
        rhs[d] -= uq[d].mean()

    def res(phi):
        L = o.stag_applyop(phi)
        return np.sqrt(sum(((rhs[d] - L[d]) ** 2).sum() for d in range(3)))

    phi = o.zeros_u()
    r = [res(phi)]
    for _ in range(8):
        phi = o.stag_mg_solve(rhs, phi)
        r.append(res(phi))
    rates = [r[i] / r[i + 1] for i in range(4, 8)]
    assert all(7.0 < q < 15.0 for q in rates), rates


def _dense(o, g):
    """Assemble the Stokes matrix column by column through apply_matrix."""
    shapes = [g.face_shape(a) for a in range(g.dm)] + [g.cc_shape()]
    sizes = [int(np.prod(s)) for s in shapes]
    N = sum(sizes)
    A = np.zeros((N, N))
    for j in range(N):
        v = np.zeros(N)
        v[j] = 1.0
        parts, o0 = [], 0
        for s, n in zip(shapes, sizes):
            parts.append(v[o0:o0 + n].reshape(s))
            o0 += n
        bu, bp = o.apply_matrix(parts[:-1], parts[-1])
        A[:, j] = np.concatenate([x.ravel() for x in bu] + [bp.ravel()])
    return A, shapes, sizes


@pytest.mark.parametrize("walls", [False, True])
def test_gmres_matches_dense_solve(walls):
    if walls:
        g = Grid((4, 4, 6), 0.25, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    else:
        g = Grid((4, 4, 4), 0.25)
    rng = np.random.default_rng(2)
    o = oracle.Oracle(g, visc_type=-1, gmres_max_inner=30, gmres_rel_tol=1e-13)
    beta = 1 + 0.5 * rng.random(g.cc_shape())
    alpha = [np.ones(g.face_shape(a)) for a in range(3)]
    bed = [1 + 0.5 * rng.random(g.edge_shape(e)) for e in range(3)]
    for e in range(3):  # duplicated periodic edges hold one value
        for d in g.edge_nodal(e):
            if g.bc_lo[d] == PERIODIC:
                ax = 2 - d
                i0 = [slice(None)] * 3
                iN = [slice(None)] * 3
                i0[ax], iN[ax] = 0, -1
                bed[e][tuple(iN)] = bed[e][tuple(i0)]
    o.set_coeffs(alpha, beta, bed, np.zeros(g.cc_shape()), 3.0)
    A, shapes, sizes = _dense(o, g)
    # a right-hand side in the range of A: b = A x_true
    x_true = rng.standard_normal(A.shape[0])
    b = A @ x_true
    parts, o0 = [], 0
    for s, n in zip(shapes, sizes):
        parts.append(b[o0:o0 + n].reshape(s))
        o0 += n
    xu, xp, st, rc = o.gmres(parts[:-1], parts[-1])
    assert rc == 0 and st.converged
    x = np.concatenate([v.ravel() for v in xu] + [xp.ravel()])
    assert np.linalg.norm(A @ x - b) <= 1e-9 * np.linalg.norm(b)
    # unique up to the pressure constant (and the dead DOFs on wall faces)
    nu = sum(sizes[:-1])
    live = np.abs(A).sum(axis=0) > 0
    xt = x_true.copy()
    xt[~live] = 0.0
    du = (x - xt)[:nu]
    dp = (x - xt)[nu:]
    assert np.abs(du[live[:nu]]).max() < 1e-8 * np.abs(xt[:nu]).max()
    assert np.abs(dp - dp.mean()).max() < 1e-8 * np.abs(xt[nu:]).max()


def test_mac_mg_converges_to_dense_solution():
    g = Grid((8, 8, 8), 1 / 8, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    rng = np.random.default_rng(4)
    alpha = [1 + rng.random(g.face_shape(a)) for a in range(3)]
    for a in range(2):
        ax = 2 - a
        i0 = [slice(None)] * 3
        iN = [slice(None)] * 3
        i0[ax], iN[ax] = 0, -1
        alpha[a][tuple(iN)] = alpha[a][tuple(i0)]
    o = oracle.Oracle(g, visc_type=1, mac_max_vcycles=60)
    one = lambda s: np.ones(s)  # noqa: E731
    o.set_coeffs(alpha, one(g.cc_shape()), [one(g.edge_shape(e)) for e in range(3)], one(g.cc_shape()), 1.0)
    r = rng.standard_normal(g.cc_shape())
    r -= r.mean()
    phi = o.mac_mg_solve(r)
    # dense D (alpha^-1 G) with zero flux through the walls
    nx, ny, nz = g.n3
    N = nx * ny * nz
    M = np.zeros((N, N))
    idx = lambda i, j, k: i + nx * (j + ny * k)  # noqa: E731
    h2 = g.dx ** 2
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                c = idx(i, j, k)
                for a, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    for s in (0, 1):  # lo face, hi face of cell c
                        fi, fj, fk = i + s * di, j + s * dj, k + s * dk
                        F = (fi, fj, fk)[a]
                        n_a = (nx, ny, nz)[a]
                        if a == 2 and (F == 0 or F == n_a):
                            continue
                        b = 1.0 / alpha[a][fk, fj, fi]
                        ni, nj, nk = (i + (di if s else -di)) % nx, (j + (dj if s else -dj)) % ny, k + (dk if s else -dk)
                        M[c, idx(ni, nj, nk)] += b / h2
                        M[c, c] -= b / h2
    x = np.linalg.lstsq(M, r.ravel(), rcond=None)[0]
    y = phi.ravel()
    assert np.abs((y - y.mean()) - (x - x.mean())).max() < 1e-9 * np.abs(x).max()


def test_stochastic_impulse_and_wall_rules():
    g = Grid((6, 6, 6), 0.5, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, SLIP_WALL))
    o = oracle.Oracle(g, visc_type=-1)
    dt, kT = 0.2, 1.3
    var = np.sqrt(2.0 * kT / (g.dx ** 3 * dt))
    eta = np.full(g.cc_shape(), 4.0)
    eed = [np.full(g.edge_shape(e), 9.0) for e in range(3)]
    zc = [np.zeros(g.cc_shape()) for _ in range(3)]
    ze = [np.zeros(g.edge_shape(q // 2)) for q in range(6)]
    # one cell-centred xx entry at cell (i,j,k) = (2,3,1): f_x = +/- var sqrt(eta)/dx on its lo / hi x-face
    W = [x.copy() for x in zc]
    W[0][1, 3, 2] = 1.0
    f = o.stochastic_m_fluxdiv(W, ze, eta, eed, dt, kT)
    v = var * 2.0 / g.dx
    assert np.isclose(f[0][1, 3, 2], v) and np.isclose(f[0][1, 3, 3], -v)
    assert np.count_nonzero(f[0]) == 2 and not np.any(f[1]) and not np.any(f[2])
    # an xz edge entry on the no-slip z-lo wall is scaled by sqrt(2); on the slip z-hi wall it is 0
    for k, fac in ((0, np.sqrt(2.0)), (6, 0.0)):
        We = [x.copy() for x in ze]
        We[2][k, 2, 3] = 1.0  # xz edge, component 1 (the x-momentum flux in z)
        f = o.stochastic_m_fluxdiv(zc, We, eta, eed, dt, kT)
        kk = 0 if k == 0 else 5
        expect = (-1.0 if k == 0 else 1.0) * fac * var * 3.0 / g.dx
        assert np.isclose(f[0][kk, 2, 3], expect, rtol=1e-14, atol=1e-300)
    # wall-normal faces carry no forcing
    rng = np.random.default_rng(0)
    Wc = [rng.standard_normal(g.cc_shape()) for _ in range(3)]
    We = [rng.standard_normal(g.edge_shape(q // 2)) for q in range(6)]
    f = o.stochastic_m_fluxdiv(Wc, We, eta, eed, dt, kT)
    assert not np.any(f[2][0]) and not np.any(f[2][-1])
    # fully periodic: each momentum forcing is a divergence and sums to zero over unique faces
    gp = Grid((6, 6, 6), 0.5)
    op = oracle.Oracle(gp, visc_type=-1)
    f = op.stochastic_m_fluxdiv(Wc, [rng.standard_normal(gp.edge_shape(q // 2)) for q in range(6)], eta,
                                [np.full(gp.edge_shape(e), 9.0) for e in range(3)], dt, kT)
    uniq = [f[0][..., :-1], f[1][:, :-1, :], f[2][:-1]]
    for a in range(3):
        assert abs(uniq[a].sum()) < 1e-12 * np.abs(uniq[a]).sum()


def test_nlevs_mg():
    from lowmachfhd_amd.grid import nlevs_mg

    for n, expect in ((128, 7), (64, 6), (48, 5), (96, 6), (192, 7), (256, 8), (32, 5)):
        g = Grid((n, n, n), 1.0 / n)
        assert nlevs_mg(g) == expect == oracle.Oracle(g).nlevs_mg(2)


@pytest.mark.parametrize("pt", [1, -1, 2, -2, 3, -3, 4, -4, 5, -5])
@pytest.mark.parametrize("vt", [-1, 2, 3])
def test_precon_types_match_dense_formulas(pt, vt):
    """apply_precon (apply_precon.f90:101-458) with its multigrid subsolves run to convergence
    equals the block formulas written with dense exact subsolves:
      A_uu^-1 (velocity block of apply_matrix), D = -(pressure row block), G (gradient block),
      L_alpha = D diag(1/alpha) G with no flux through walls, Schur scaling chi = beta, 2 beta,
      4/3 beta + gamma; the pressure mean removed at the end."""
    g = Grid((4, 4, 6), 0.25, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    rng = np.random.default_rng(7)
    theta = 3.0
    kn = dict(stag_max_vcycles=400, stag_rel_tol=1e-15, mac_max_vcycles=400, precon_type=pt)
    o = oracle.Oracle(g, visc_type=vt, **kn)
    beta = 1 + 0.5 * rng.random(g.cc_shape())
    gamma = 0.3 * rng.random(g.cc_shape())
    alpha = [1 + 0.5 * rng.random(g.face_shape(a)) for a in range(3)]
    bed = [1 + 0.5 * rng.random(g.edge_shape(e)) for e in range(3)]
    for arrs, nodal in [(alpha, lambda a: [a]), (bed, g.edge_nodal)]:
        for q, x in enumerate(arrs):
            for d in nodal(q):
                if g.bc_lo[d] == PERIODIC:
                    ax = 2 - d
                    i0 = [slice(None)] * 3
                    iN = [slice(None)] * 3
                    i0[ax], iN[ax] = 0, -1
                    x[tuple(iN)] = x[tuple(i0)]
    if abs(vt) != 3:
        gamma[:] = 0.0 if vt < 0 else gamma
    o.set_coeffs(alpha, beta, bed, gamma, theta)
    A, shapes, sizes = _dense(o, g)
    nu = sum(sizes[:-1])
    Auu, G, Dm = A[:nu, :nu], A[:nu, nu:], -A[nu:, :nu]
    live = np.abs(Auu).sum(axis=0) > 0  # wall faces are fixed at zero
    def Ainv(r):
        x = np.zeros(nu)
        x[live] = np.linalg.solve(Auu[np.ix_(live, live)], r[live])
        return x
    a_inv = np.concatenate([1.0 / x.ravel() for x in alpha])
    a_inv[~live] = 0.0
    L = Dm @ (a_inv[:, None] * G)
    Linv = lambda r: np.linalg.lstsq(L, r, rcond=None)[0]  # noqa: E731
    av = abs(vt)
    b_, g_ = beta.ravel(), gamma.ravel()
    chi = b_ if av == 1 else (2 * b_ if av == 2 else 4 / 3 * b_ + g_)
    # vectors of the solution space, as every preconditioner input inside GMRES is: duplicated
    # periodic faces equal, and a mean-free pressure part (it is -D of a velocity), so that
    # the MAC right-hand side is consistent; an inconsistent one is outside what FBoxLib's
    # solver (absent here) would pin
    bu = [rng.standard_normal(s) for s in shapes[:-1]]
    for a in range(2):
        ax = 2 - a
        i0 = [slice(None)] * 3
        iN = [slice(None)] * 3
        i0[ax], iN[ax] = 0, -1
        bu[a][tuple(iN)] = bu[a][tuple(i0)]
    bp = rng.standard_normal(shapes[-1])
    bp -= bp.mean()
    Bu = np.concatenate([x.ravel() for x in bu])
    Bu[~live] = 0.0
    Bp = bp.ravel()
    apt = abs(pt)
    if apt == 1:
        xs = Ainv(Bu)
        r = Dm @ xs + Bp
        phi = Linv(r)
        xu = xs - a_inv * (G @ phi)
        xp = -chi * (r if pt == 1 else L @ phi) + theta * phi
    elif apt in (2, 5):
        xu = Ainv(Bu)
        r = Dm @ xu + Bp
        phi = Linv(r)
        xp = -chi * r + theta * phi
        if pt < 0:
            xp = -xp
        if apt == 5:
            xu = Ainv(Bu - G @ xp)
    elif apt == 3:
        r = -Bp if pt < 0 else Bp
        phi = Linv(r)
        xp = -chi * r + theta * phi
        xu = Ainv(Bu - G @ xp)
    else:
        xu = Ainv(Bu)
        phi = Linv(Bp)
        v = b_ * Bp - theta * phi
        c = {(1, 4): -1.0, (2, 4): -2.0, (2, -4): 2.0, (3, 4): -4 / 3, (3, -4): 4 / 3}.get((av, pt), 1.0)
        xp = v * c + ({4: -1.0, -4: 1.0}[pt] * g_ * Bp if av == 3 else 0.0)
    xp = xp - xp.mean()
    ou, op = o.apply_precon(bu, bp)
    OU = np.concatenate([x.ravel() for x in ou])
    assert np.abs(OU - xu)[live].max() < 1e-9 * np.abs(xu).max(), (pt, vt)
    assert np.abs(op.ravel() - xp).max() < 1e-9 * np.abs(xp).max(), (pt, vt)


@pytest.mark.parametrize("pt", [1, -1, 2, -3, 4, 5])
@pytest.mark.parametrize("vt", [-1, -2, 2])
def test_visc_schur_matches_dense_formulas(pt, vt):
    """visc_schur_approx = 1 (visc_schur_complement, apply_precon.f90:538-685) with the
    multigrid subsolves run to convergence equals, on a periodic box, the dense formula
      x_p = L_mu^+ D Minv (-A0) Minv G L_mu^+ r + theta phi,   L_mu = D Minv G,
    where A0 is the velocity block of the theta = 0 matrix and Minv = 1 / (h^2 diag A0) read
    off that dense matrix (an independent check of inverse_diag_lap); r is the MAC right-hand
    side of each precon_type (D x_u* + b_p for 1, 2, 5; +-b_p for 3 and 4) and phi its
    L_alpha solve; the rest of each type as in the simple-Schur test above."""
    g = Grid((4, 4, 6), 0.25)
    rng = np.random.default_rng(11)
    theta = 3.0
    kn = dict(stag_max_vcycles=400, stag_rel_tol=1e-15, mac_max_vcycles=400, precon_type=pt, visc_schur_approx=1)
    beta = (1 + 0.5 * rng.random(g.cc_shape())) if vt < 0 else np.full(g.cc_shape(), 1.3)
    alpha = [1 + 0.5 * rng.random(g.face_shape(a)) for a in range(3)]
    bed = [(1 + 0.5 * rng.random(g.edge_shape(e))) if vt < 0 else np.full(g.edge_shape(e), 1.3) for e in range(3)]
    for arrs, nodal in [(alpha, lambda a: [a]), (bed, g.edge_nodal)]:
        for q, x in enumerate(arrs):
            for d in nodal(q):
                ax = 2 - d
                i0 = [slice(None)] * 3
                iN = [slice(None)] * 3
                i0[ax], iN[ax] = 0, -1
                x[tuple(iN)] = x[tuple(i0)]
    gamma = np.zeros(g.cc_shape())
    o = oracle.Oracle(g, visc_type=vt, **kn)
    o.set_coeffs(alpha, beta, bed, gamma, theta)
    o0 = oracle.Oracle(g, visc_type=vt)
    o0.set_coeffs(alpha, beta, bed, gamma, 0.0)
    A, shapes, sizes = _dense(o, g)
    A0, _, _ = _dense(o0, g)
    nu = sum(sizes[:-1])
    Auu, G, Dm = A[:nu, :nu], A[:nu, nu:], -A[nu:, :nu]
    negA0 = A0[:nu, :nu]                       # (0 alpha - A0)
    minv = 1.0 / (g.dx ** 2 * np.diag(-negA0))  # 1 / (h^2 diag A0)
    Ainv = lambda r: np.linalg.solve(Auu, r)  # noqa: E731
    a_inv = np.concatenate([1.0 / x.ravel() for x in alpha])
    L = Dm @ (a_inv[:, None] * G)
    Lmu = Dm @ (minv[:, None] * G)
    Linv = lambda M, r: np.linalg.lstsq(M, r, rcond=None)[0]  # noqa: E731

    def visc(r, phi):
        pm = Linv(Lmu, r)
        r2 = Dm @ (minv * (negA0 @ (minv * (G @ pm))))
        return Linv(Lmu, r2) + theta * phi

    bu = [rng.standard_normal(s) for s in shapes[:-1]]
    for a in range(3):
        ax = 2 - a
        i0 = [slice(None)] * 3
        iN = [slice(None)] * 3
        i0[ax], iN[ax] = 0, -1
        bu[a][tuple(iN)] = bu[a][tuple(i0)]
    bp = rng.standard_normal(shapes[-1])
    bp -= bp.mean()
    Bu = np.concatenate([x.ravel() for x in bu])
    Bp = bp.ravel()
    apt = abs(pt)
    if apt == 1:
        xs = Ainv(Bu)
        r = Dm @ xs + Bp
        phi = Linv(L, r)
        xu = xs - a_inv * (G @ phi)
        xp = visc(r, phi)
    elif apt in (2, 5):
        xu = Ainv(Bu)
        r = Dm @ xu + Bp
        xp = visc(r, Linv(L, r))
        if pt < 0:
            xp = -xp
        if apt == 5:
            xu = Ainv(Bu - G @ xp)
    elif apt == 3:
        r = -Bp if pt < 0 else Bp
        xp = visc(r, Linv(L, r))
        xu = Ainv(Bu - G @ xp)
    else:
        xu = Ainv(Bu)
        xp = visc(Bp, Linv(L, Bp))
    xp = xp - xp.mean()
    ou, op = o.apply_precon(bu, bp)
    OU = np.concatenate([x.ravel() for x in ou])
    assert np.abs(OU - xu).max() < 1e-9 * np.abs(xu).max(), (pt, vt)
    assert np.abs(op.ravel() - xp).max() < 1e-9 * np.abs(xp).max(), (pt, vt)
