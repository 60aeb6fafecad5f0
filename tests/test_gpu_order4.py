"""gmres_spatial_order = 4 on the GPU (apply_matrix with stag_applyop_4o + compute_grad_4o,
apply_matrix.f90:123-135) against the oracle on the same inputs: apply_matrix within 1e-13
relative, GMRES iteration counts within +-1 and solutions within 1e-10, periodic 2-D and
3-D, FP64 and mixed forms."""
import numpy as np
import pytest

from lowmachfhd_amd.grid import Grid

pytestmark = pytest.mark.gpu


def rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


def _setup(dims, mixed=False):
    from oracle import Oracle

    from lowmachfhd_amd import solver as S

    g = Grid(dims, 1.0 / dims[0])
    kn = dict(gmres_max_inner=30, gmres_rel_tol=1e-11, gmres_spatial_order=4)
    o = Oracle(g, visc_type=1, **kn)
    s = S.StokesSolver(g, visc_type=1, precision=S.MIXED if mixed else S.FP64,
                       krylov=S.FGMRES if mixed else S.GMRES_LEFT, **kn)
    args = ([np.ones(g.face_shape(a)) for a in range(g.dm)], 0.8 * np.ones(g.cc_shape()),
            [0.8 * np.ones(g.edge_shape(e)) for e in range(g.nedge)], np.zeros(g.cc_shape()), 20.0)
    o.set_coeffs(*args)
    s.set_coeffs(*args)
    return g, o, s


@pytest.mark.parametrize("dims", [(32, 24, 16), (48, 40)])
def test_apply_matrix_order4(dims):
    g, o, s = _setup(dims)
    rng = np.random.default_rng(7)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    p = rng.standard_normal(g.cc_shape())
    b1, b2 = o.apply_matrix(u, p), s.apply_matrix(u, p)
    assert rel(b2[0], b1[0]) < 1e-13 and rel([b2[1]], [b1[1]]) < 1e-13
    s.close()


@pytest.mark.parametrize("mixed", [0, 1])
@pytest.mark.parametrize("dims", [(32, 24, 16), (48, 40)])
def test_gmres_order4(dims, mixed):
    g, o, s = _setup(dims, bool(mixed))
    rng = np.random.default_rng(8)
    bu = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    for a in range(g.dm):  # periodic data: the duplicated face n holds face 0's value (else the
        np.moveaxis(bu[a], 2 - a, 0)[-1] = np.moveaxis(bu[a], 2 - a, 0)[0]  # system is inconsistent)
    bp = np.zeros(g.cc_shape())
    xu1, xp1, st1, rc = o.gmres(bu, bp)
    xu2, xp2, st2 = s.gmres(bu, bp)
    eu, ep = rel(xu2, xu1), rel([xp2 - xp2.mean()], [xp1 - xp1.mean()])
    print(f"{dims} mixed={mixed}: iterations {st1.total_iter}/{st2.iters} u {eu:.1e} p {ep:.1e}", flush=True)
    assert rc == 0 and st1.converged == 1 and st2.converged == 1
    if mixed:
        assert st2.true_rel_resid <= 1e-11 and eu < 1e-8 and ep < 1e-8
    else:
        assert abs(st1.total_iter - st2.iters) <= 1 and eu < 1e-10 and ep < 1e-10
    s.close()

