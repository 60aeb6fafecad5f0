"""gmres_spatial_order = 4 in the oracle (stag_applyop_4o + compute_grad_4o inside apply_matrix,
apply_matrix.f90:123-135), pinned on the CPU:
  * stag_applyop_4o (visc_type 1) converges at fourth order to (theta alpha - beta Laplacian) u
    for smooth periodic fields, 2-D and 3-D -- any wrong weight or offset drops the order;
  * compute_grad_4o is pinned by its Fourier symbol: on p = cos(k x + phi) the reference's
    weights (-1, 15, -15, 1)/(12 h) give exactly sin(k x_f + phi) (2 sin(3kh/2) - 30 sin(kh/2))
    / (12 h) on every x-face (they are a second-order formula: the fourth-order staggered
    weights would be (-1, 27, -27, 1)/24, which this check tells apart);
  * GMRES on the order-4 system (second-order preconditioner, as in the reference) converges,
    and its solution satisfies the order-4 equations."""
import numpy as np
import pytest

from lowmachfhd_amd.grid import Grid

TP = 2 * np.pi


def _oracle(g, **kn):
    from oracle import Oracle

    o = Oracle(g, visc_type=1, gmres_spatial_order=4, **kn)
    return o


def _coef(o, g, b=0.7, theta=3.0, alpha=1.3):
    o.set_coeffs([np.full(g.face_shape(a), alpha) for a in range(g.dm)], b * np.ones(g.cc_shape()),
                 [b * np.ones(g.edge_shape(e)) for e in range(g.nedge)], np.zeros(g.cc_shape()), theta)


@pytest.mark.parametrize("dm", [2, 3])
def test_applyop_4o_fourth_order(dm):
    b, theta, alpha = 0.7, 3.0, 1.3
    errs = []
    for n in ((16, 32, 64) if dm == 2 else (8, 16, 32)):
        g = Grid((n,) * dm, 1.0 / n)
        o = _oracle(g)
        _coef(o, g, b, theta, alpha)
        u, ex = [], []
        for a in range(dm):
            c = g.face_coords(a)
            f = np.sin(TP * c[0] + 0.3 * a) * np.cos(TP * c[1]) * (np.sin(TP * c[2]) if dm == 3 else 1.0)
            u.append(f)
            ex.append(theta * alpha * f + b * dm * TP ** 2 * f)
        bu, _ = o.apply_matrix(u, np.zeros(g.cc_shape()))
        errs.append(max(np.abs(bu[a] - ex[a]).max() for a in range(dm)))
    print(dm, errs)
    assert errs[0] / errs[1] > 14 and errs[1] / errs[2] > 15


@pytest.mark.parametrize("k", [1, 3])
def test_grad_4o_fourier_symbol(k):
    n = 16
    g = Grid((n, n, n), 1.0 / n)
    o = _oracle(g)
    _coef(o, g)
    h = 1.0 / n
    kap = TP * k
    Xc = g.cc_coords()[0]
    p = np.cos(kap * Xc + 0.4)
    bu, _ = o.apply_matrix(o.zeros_u(), p)
    Xf = g.face_coords(0)[0]
    expect = np.sin(kap * Xf + 0.4) * (2 * np.sin(1.5 * kap * h) - 30 * np.sin(0.5 * kap * h)) / (12 * h)
    assert np.abs(bu[0] - expect).max() < 1e-11 * np.abs(expect).max()
    assert np.abs(bu[1]).max() < 1e-11 * np.abs(expect).max() and np.abs(bu[2]).max() < 1e-11 * np.abs(expect).max()


@pytest.mark.parametrize("dm", [2, 3])
def test_gmres_order4_solves_order4_system(dm):
    n = 16
    g = Grid((n,) * dm, 1.0 / n)
    o = _oracle(g, gmres_max_inner=30, gmres_rel_tol=1e-11)
    _coef(o, g, b=1.0, theta=10.0, alpha=1.0)
    rng = np.random.default_rng(3)
    bu = [rng.standard_normal(g.face_shape(a)) for a in range(dm)]
    for a in range(dm):  # periodic data: face n holds face 0's value
        np.moveaxis(bu[a], 2 - a, 0)[-1] = np.moveaxis(bu[a], 2 - a, 0)[0]
    bp = np.zeros(g.cc_shape())
    xu, xp, st, rc = o.gmres(bu, bp)
    assert rc == 0 and st.converged == 1
    ru, rp = o.apply_matrix(xu, xp)
    scale = max(np.abs(x).max() for x in bu)
    assert max(np.abs(ru[a] - bu[a]).max() for a in range(dm)) < 1e-8 * scale
    assert np.abs(rp).max() < 1e-8 * scale / g.dx
