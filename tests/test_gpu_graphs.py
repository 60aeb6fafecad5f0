"""GMRES Arnoldi steps and preconditioner applications replayed as HIP graphs (lmfhd_desc.graphs
= 1) give the same bits as the launch-by-launch path, on every replay, with eager work and
synchronisations between the replays."""
import numpy as np
import pytest

from lowmachfhd_amd import problems
from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, Grid

pytestmark = pytest.mark.gpu


def _cases():
    per = Grid((32, 32, 32), 1 / 32)
    wall = Grid((24, 24, 24), 1 / 24, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    return [("per32", per, problems.config1_like, 7), ("zwall24", wall, problems.config2_like, 8)]


def _solver(g, mixed, graphs=0):
    from lowmachfhd_amd import solver as S

    return S.StokesSolver(g, visc_type=-1, precision=S.MIXED if mixed else S.FP64,
                          krylov=S.FGMRES if mixed else S.GMRES_LEFT, gmres_max_inner=30, graphs=graphs)


@pytest.mark.parametrize("mixed", [0, 1])
@pytest.mark.parametrize("case", [0, 1])
def test_graph_replays_match_eager(case, mixed):
    import torch

    name, g, recipe, seed = _cases()[case]
    e = _solver(g, mixed, graphs=0)
    prob = recipe(g, seed=seed, averager=e)
    args = (prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
    e.set_coeffs(*args)
    b = torch.from_numpy(e.pack(prob["b_u"], prob["b_p"])).to("cuda:0")
    x = torch.zeros_like(b)
    torch.cuda.synchronize()
    st = e.gmres_dev(b.data_ptr(), x.data_ptr())
    e.sync()
    assert st.converged == 1
    ref = x.cpu().numpy()
    # the preconditioner alone, too
    pe = torch.zeros_like(b)
    e.apply_precon_dev(b.data_ptr(), pe.data_ptr())
    e.sync()
    pref = pe.cpu().numpy()
    e.close()

    s = _solver(g, mixed, graphs=1)
    s.set_coeffs(*args)
    for rep in range(3):
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        st2 = s.gmres_dev(b.data_ptr(), x.data_ptr())
        s.sync()
        _ = float((b * 2.0).sum())  # eager torch work between the replays
        assert st2.iters == st.iters
        d = np.abs(x.cpu().numpy() - ref).max()
        print(f"{name} mixed={mixed} replay {rep}: max |x_graph - x_eager| = {d:.3e}", flush=True)
        assert d == 0.0
        p = torch.zeros_like(b)
        s.apply_precon_dev(b.data_ptr(), p.data_ptr())
        s.sync()
        assert np.array_equal(p.cpu().numpy(), pref)
    # new coefficients drop the graphs: the next solve captures again and still matches
    s.set_coeffs(*args)
    x = torch.zeros_like(b)
    s.gmres_dev(b.data_ptr(), x.data_ptr())
    s.sync()
    assert np.array_equal(x.cpu().numpy(), ref)
    s.close()
