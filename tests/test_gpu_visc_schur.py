"""visc_schur_approx = 1 on the GPU (the pressure block of the preconditioner adds
visc_schur_complement, apply_precon.f90:538-600: a MAC solve with face coefficient 1/diag(L),
the viscous operator at theta 0 weighted by 1/diag(L) on both sides, a second MAC solve)
against the oracle on the same inputs: one preconditioner application within 1e-12 relative,
GMRES iteration counts within +-1 and solutions within 1e-10 (FP64), and FGMRES to the true
relative residual 1e-10 in the mixed form. Periodic, no-slip and slip walls, 2-D and 3-D,
visc_type -1 and 2, every precon_type."""
import numpy as np
import pytest
from test_gpu_parity import consistent_u, random_coeffs

from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid

pytestmark = pytest.mark.gpu

P, W, SW = PERIODIC, NO_SLIP_WALL, SLIP_WALL
GRIDS = {
    "per3": lambda: Grid((32, 24, 16), 1 / 32),
    "zwall3": lambda: Grid((32, 24, 16), 1 / 32, bc_lo=(P, P, W), bc_hi=(P, P, W)),
    "xwall2": lambda: Grid((48, 40), 1 / 48, bc_lo=(W, P), bc_hi=(W, P)),
    "yslip3": lambda: Grid((24, 24, 16), 1 / 24, bc_lo=(P, SW, P), bc_hi=(P, SW, P)),
}


def rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


def _setup(name, vt, pt, mixed=False):
    from oracle import Oracle

    from lowmachfhd_amd import solver as S

    g = GRIDS[name]()
    rng = np.random.default_rng(abs(vt) * 10 + pt + 3)
    kn = dict(gmres_max_inner=30, gmres_rel_tol=1e-11, precon_type=pt, visc_schur_approx=1)
    o = Oracle(g, visc_type=vt, **kn)
    s = S.StokesSolver(g, visc_type=vt, precision=S.MIXED if mixed else S.FP64,
                       krylov=S.FGMRES if mixed else S.GMRES_LEFT, **kn)
    alpha, beta, bed, gamma = random_coeffs(g, rng, vt < 0)
    args = (alpha, beta, bed, gamma, 10.0)
    o.set_coeffs(*args)
    s.set_coeffs(*args)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = rng.standard_normal(g.cc_shape())
    return g, o, s, bu, bp


@pytest.mark.parametrize("pt", [1, -1, 2, -3, 4, 5])
@pytest.mark.parametrize("vt", [-1, 2])
@pytest.mark.parametrize("name", list(GRIDS))
def test_visc_schur_fp64(name, vt, pt):
    g, o, s, bu, bp = _setup(name, vt, pt)
    xu1, xp1 = o.apply_precon(bu, bp)
    xu2, xp2 = s.apply_precon(bu, bp)
    assert rel(xu2, xu1) < 1e-12 and rel([xp2], [xp1]) < 1e-12
    yu1, yp1, st1, rc = o.gmres(bu, 0 * bp)
    yu2, yp2, st2 = s.gmres(bu, 0 * bp)
    eu = rel(yu2, yu1)
    print(f"{name} vt={vt} pt={pt}: iterations {st1.total_iter}/{st2.iters} u {eu:.1e}", flush=True)
    assert rc == 0 and st1.converged == 1 and st2.converged == 1 and st2.true_rel_resid < 1e-8
    assert abs(st1.total_iter - st2.iters) <= 1 and eu < 1e-10
    s.close()


@pytest.mark.parametrize("pt", [1, -3])
@pytest.mark.parametrize("name", ["per3", "zwall3"])
def test_visc_schur_mixed(name, pt):
    g, o, s, bu, bp = _setup(name, -1, pt, mixed=True)
    yu1, yp1, st1, rc = o.gmres(bu, 0 * bp)
    yu2, yp2, st2 = s.gmres(bu, 0 * bp)
    eu = rel(yu2, yu1)
    print(f"{name} mixed pt={pt}: iterations {st1.total_iter}/{st2.iters} u {eu:.1e} "
          f"true rel resid {st2.true_rel_resid:.1e}", flush=True)
    assert st2.converged == 1 and st2.true_rel_resid <= 1e-10 and eu < 1e-8
    s.close()
