"""The oracle's restatement of filter_stoch_m_flux (stochastic_m_fluxdiv.f90:604-778), pinned
through stochastic_m_fluxdiv by what the filters' design fixes, not by their taps: each pass is
a symmetric circular convolution along one axis, so a stress that is one Fourier mode along an
axis and constant along the others comes out multiplied by the filter's response there (the
constant directions see the DC gain, which must be exactly 1). The three filters are the
maximally flat ones of their widths: the Nyquist mode (a checkerboard) is annihilated, and a
smooth mode of angle theta keeps 1 - O(theta^2) (width 1), 1 - O(theta^4) (width 2) and
1 - O(theta^8) (width 4). A tap typed wrong, a pass left out or a pass along the wrong axis
breaks one of these."""
import numpy as np
import pytest

from lowmachfhd_amd.grid import Grid

ORDER = {1: 2, 2: 4, 4: 8}


def _fluxdiv(g, w, W_cc, W_ed):
    import oracle

    o = oracle.Oracle(g, visc_type=-1, filtering_width=w)
    eta = np.ones(g.cc_shape())
    eta_ed = [np.ones(g.edge_shape(e)) for e in range(g.nedge)]
    return o.stochastic_m_fluxdiv(W_cc, W_ed, eta, eta_ed, dt=1.0, kT=0.5)


def _mode(shape, axis_np, m, n, nodal):
    idx = np.arange(shape[axis_np])
    v = np.cos(2 * np.pi * m * idx / n)  # on a nodal axis index n repeats index 0
    sh = [1, 1, 1]
    sh[axis_np] = shape[axis_np]
    return np.broadcast_to(v.reshape(sh), shape).copy()


def _response(g, w, kind, d, m):
    """ratio filtered / unfiltered of the flux divergence for a stress that is the mode
    cos(2 pi m i / n) along axis d: kind 'cc' puts it in the cell stress of component d,
    kind 'ed' in an edge stress that component d reads along a transverse direction."""
    n = g.n3[d]
    W_cc = [np.zeros(g.cc_shape()) for _ in range(g.dm)]
    W_ed = [np.zeros(g.edge_shape(e)) for e in range(g.nedge) for _ in range(2)]
    if kind == "cc":
        W_cc[d][...] = _mode(g.cc_shape(), 2 - d, m, n, False)
        comp = d
    else:  # edge stress of the first edge array nodal along d, component 0
        e = next(e for e in range(g.nedge) if d in g.edge_nodal(e))
        W_ed[2 * e][...] = _mode(g.edge_shape(e), 2 - d, m, n, True)
        nd = g.edge_nodal(e)
        comp = min(nd) if g.dm == 3 else 0  # component 0 of an edge stress feeds the lower direction
        if comp == d:
            comp = max(nd)
        W_ed[2 * e + 1][...] = W_ed[2 * e]
    f0 = _fluxdiv(g, 0, W_cc, W_ed)
    fw = _fluxdiv(g, w, W_cc, W_ed)
    a = [q for q in range(g.dm) if np.abs(f0[q]).max() > 0]
    num = sum(float(np.vdot(f0[q], fw[q])) for q in a)
    den = sum(float(np.vdot(f0[q], f0[q])) for q in a)
    ratio = num / den
    resid = max(np.abs(fw[q] - ratio * f0[q]).max() for q in a) / max(np.abs(f0[q]).max() for q in a)
    return ratio, resid


GRIDS = {2: lambda: Grid((32, 32), 1 / 32), 3: lambda: Grid((32, 32, 32), 1 / 32)}


@pytest.mark.parametrize("kind", ["cc", "ed"])
@pytest.mark.parametrize("w", [1, 2, 4])
@pytest.mark.parametrize("dm", [2, 3])
def test_filter_nyquist_and_flatness(dm, w, kind):
    g = GRIDS[dm]()
    for d in range(dm):
        r_ny, res = _response(g, w, kind, d, 16)
        assert abs(r_ny) < 1e-13 and res < 1e-13, (d, r_ny, res)
        r1, res1 = _response(g, w, kind, d, 2)
        r2, res2 = _response(g, w, kind, d, 1)
        assert res1 < 1e-12 and res2 < 1e-12  # each mode is an eigenvector: pure scaling
        assert 0 < 1 - r1 < 0.05 and 0 < 1 - r2 < 1 - r1
        p = np.log2((1 - r1) / (1 - r2))
        assert abs(p - ORDER[w]) < 0.15, (d, p)


@pytest.mark.parametrize("dm", [2, 3])
def test_filter_zero_is_identity_and_linear(dm):
    g = GRIDS[dm]()
    rng = np.random.default_rng(3)
    W_cc = [rng.standard_normal(g.cc_shape()) for _ in range(g.dm)]
    W_ed = [rng.standard_normal(g.edge_shape(e)) for e in range(g.nedge) for _ in range(2)]
    f1 = _fluxdiv(g, 2, W_cc, W_ed)
    f2 = _fluxdiv(g, 2, [2 * x for x in W_cc], [2 * x for x in W_ed])
    assert max(np.abs(2 * a - b).max() for a, b in zip(f1, f2)) < 1e-12 * max(np.abs(a).max() for a in f1)
    # white noise loses variance: sum of squared taps per pass, (3/8)^dm for width 1 on the
    # stress; the divergence's variance drops by a comparable factor
    f0 = _fluxdiv(g, 0, W_cc, W_ed)
    fw = _fluxdiv(g, 1, W_cc, W_ed)
    assert sum(float(np.vdot(x, x)) for x in fw) < 0.5 * sum(float(np.vdot(x, x)) for x in f0)
