"""Lock-step batches of preconditioner applications (lmfhd_desc.batch = 1): solves of one size
class running at the same time share one launch sequence for their preconditioners (the k-th
launch of every solve in one launch, one grid slice per solve).  Each result must be the bits of
the same solve run alone, with the same iteration count, for the FP64 and the mixed forms, with
two size classes in flight at once."""
import numpy as np
import pytest

from lowmachfhd_amd import problems
from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, Grid

pytestmark = pytest.mark.gpu


def _items(S, grids_seeds, prec, kry, batch):
    import torch

    out = []
    for g, seed in grids_seeds:
        s = S.StokesSolver(g, visc_type=-1, precision=prec, krylov=kry, gmres_max_inner=30, batch=batch)
        prob = problems.config1_like(g, seed=seed, averager=s) if g.bc_lo[2] == PERIODIC else \
            problems.config2_like(g, seed=seed, averager=s)
        s.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
        b = torch.from_numpy(s.pack(prob["b_u"], prob["b_p"])).to("cuda:0")
        out.append((s, b, torch.zeros_like(b)))
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_equals_alone(mixed):
    import torch

    from lowmachfhd_amd import solver as S

    prec, kry = (S.MIXED, S.FGMRES) if mixed else (S.FP64, S.GMRES_LEFT)
    zw = lambda n: Grid((n, n, n), 1 / n, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL),  # noqa: E731
                        bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    gs = [(Grid((32, 32, 32), 1 / 32), 100 + q) for q in range(5)] + [(zw(24), 200 + q) for q in range(3)]
    # alone: unbatched contexts, one solve at a time
    ref = []
    for s, b, x in _items(S, gs, prec, kry, 0):
        st = s.gmres_dev(b.data_ptr(), x.data_ptr())
        s.sync()
        assert st.converged == 1
        ref.append((x.clone(), st.iters))
        s.close()
    # batched: all eight at once (two size classes), then again (the tables are reused)
    items = _items(S, gs, prec, kry, 1)
    for rep in range(2):
        for _, _, x in items:
            x.zero_()
        torch.cuda.synchronize()
        stats = S.gmres_dev_batch([it[0] for it in items], [it[1].data_ptr() for it in items],
                                  [it[2].data_ptr() for it in items], max_threads=len(items))
        for (s, b, x), (xr, ir), st in zip(items, ref, stats):
            assert st.converged == 1 and st.iters == ir, (rep, st.iters, ir)
            assert torch.equal(x, xr), rep
    # one batched context alone is its own round
    s, b, x = items[0]
    x.zero_()
    torch.cuda.synchronize()
    st = s.gmres_dev(b.data_ptr(), x.data_ptr())
    s.sync()
    assert st.iters == ref[0][1] and torch.equal(x, ref[0][0])
    for s, _, _ in items:
        s.close()
