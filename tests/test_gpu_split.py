"""Multigrid levels in parity-split storage (every extent >= 32, |visc_type| 1, Gauss-Seidel): the
staggered multigrid, the preconditioners, GMRES and the multigrid-precision operator entry
against the oracle on grids whose finest level is split, with walls of each kind.  The split
layout changes where values live, not the arithmetic: the oracle tolerances are the natural
layout's (stag_mg_solver.f90:360-398 for the sweep order)."""
import numpy as np
import pytest

from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid
from test_gpu_parity import _solver_mod, consistent_u, pair, rel

pytestmark = pytest.mark.gpu

GRIDS = {
    "per32": lambda: Grid((32, 32, 32), 1 / 32),
    # y slip walls and z no-slip walls on a split finest level (x faces: odd extent 33)
    "walls": lambda: Grid((32, 40, 36), 1 / 32, bc_lo=(PERIODIC, SLIP_WALL, NO_SLIP_WALL),
                          bc_hi=(PERIODIC, SLIP_WALL, NO_SLIP_WALL)),
    "xwall": lambda: Grid((34, 32, 32), 1 / 32, bc_lo=(NO_SLIP_WALL, PERIODIC, PERIODIC),
                          bc_hi=(NO_SLIP_WALL, PERIODIC, PERIODIC)),
    "wall2": lambda: Grid((64, 36), 1 / 64, bc_lo=(PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, NO_SLIP_WALL)),
}


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("vt", [1, -1])
def test_split_stag_mg_solver(gname, vt):
    S = _solver_mod()
    g = GRIDS[gname]()
    rng = np.random.default_rng(7 + vt)
    o, s = pair(g, vt, 5.0, rng)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    ref = o.stag_mg_solve(bu)
    assert rel(s.stag_mg_solver(bu), ref) < 1e-12
    s.close()
    sm = S.StokesSolver(g, visc_type=vt, precision=S.MIXED, krylov=S.FGMRES)
    sm.set_coeffs(*o.coeffs)
    assert rel(sm.stag_mg_solver(bu), ref) < 2e-6  # FP32 multigrid
    sm.close()


@pytest.mark.parametrize("gname", ["walls", "wall2"])
@pytest.mark.parametrize("pt", [1, -1, 2, -2, 3, 4, 5, -5])
def test_split_apply_precon(gname, pt):
    g = GRIDS[gname]()
    rng = np.random.default_rng(20 + pt)
    o, s = pair(g, -1, 5.0, rng, precon_type=pt)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = rng.standard_normal(g.cc_shape())
    bp -= bp.mean()
    xu, xp = s.apply_precon(bu, bp)
    ru, rp = o.apply_precon(bu, bp)
    assert rel(xu, ru) < 1e-11 and rel([xp], [rp]) < 1e-11


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("vt", [1, -1])
def test_split_gmres(gname, vt):
    g = GRIDS[gname]()
    rng = np.random.default_rng(31 + vt)
    o, s = pair(g, vt, 5.0, rng, gmres_max_inner=30)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = np.zeros(g.cc_shape())
    xu1, xp1, st1, rc = o.gmres(bu, bp)
    xu2, xp2, st2 = s.gmres(bu, bp)
    assert rc == 0 and st2.converged == 1
    assert abs(st2.iters - st1.total_iter) <= 1
    assert rel(xu2, xu1) < 1e-10 and rel(xp2, xp1) < 1e-10


@pytest.mark.parametrize("gname", list(GRIDS))
def test_split_applyop_mg_entry(gname):
    """lmfhd_stag_applyop_mg_dev takes and returns natural storage whatever the level's layout."""
    import torch

    S = _solver_mod()
    g = GRIDS[gname]()
    rng = np.random.default_rng(3)
    o, s = pair(g, -1, 5.0, rng)
    phi = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    ref = o.stag_applyop(phi, 0)
    vec = torch.from_numpy(s.pack(phi, s.zeros_p())).to("cuda:0")
    out = torch.zeros_like(vec)
    s.stag_applyop_mg_dev(vec.data_ptr(), out.data_ptr())
    s.sync()
    got, _ = s.unpack(out.cpu().numpy())
    assert rel(got, ref) < 1e-13
    s.close()
