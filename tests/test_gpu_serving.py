"""Concurrent solves on independent contexts (serving / ensembles).

Each result must be bit-for-bit the same when solved alone and when solved while other
contexts run on the device: every reduction in the library has a fixed order and contexts
share no state.
"""
import numpy as np
import pytest

from lowmachfhd_amd import problems
from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, Grid

pytestmark = pytest.mark.gpu


def _setup(grids, precision, krylov, seed0=5):
    import torch

    from lowmachfhd_amd import solver as S

    out = []
    for q, g in enumerate(grids):
        s = S.StokesSolver(g, visc_type=-1, precision=precision, krylov=krylov, gmres_max_inner=30)
        prob = problems.config1_like(g, seed=seed0 + q, averager=s)
        s.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
        b = torch.from_numpy(s.pack(prob["b_u"], prob["b_p"])).to("cuda:0")
        x = torch.zeros_like(b)
        out.append((s, b, x))
    return out


GRIDS = [
    lambda: Grid((32, 32, 32), 1 / 32),
    lambda: Grid((24, 24, 24), 1 / 24, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL)),
    lambda: Grid((48, 48, 48), 1 / 48),
    lambda: Grid((64, 32), 1 / 64),
]


@pytest.mark.parametrize("mixed", [False, True])
def test_alone_equals_in_the_mix(mixed):
    import torch

    from lowmachfhd_amd import solver as S

    prec, kry = (S.MIXED, S.FGMRES) if mixed else (S.FP64, S.GMRES_LEFT)
    items = _setup([f() for f in GRIDS], prec, kry)
    alone = []
    for s, b, x in items:
        x.zero_()
        torch.cuda.synchronize()
        st = s.gmres_dev(b.data_ptr(), x.data_ptr())
        s.sync()
        assert st.converged == 1
        alone.append((x.clone(), st.iters))
    # all at once through the batch entry point
    for s, b, x in items:
        x.fill_(0.0)
    torch.cuda.synchronize()
    stats = S.gmres_dev_batch([it[0] for it in items], [it[1].data_ptr() for it in items],
                              [it[2].data_ptr() for it in items], max_threads=4)
    for (s, b, x), (xa, ia), st in zip(items, alone, stats):
        assert st.iters == ia
        assert torch.equal(x, xa)
    # asynchronous submission, staggered, polled
    for s, b, x in items:
        x.fill_(0.0)
    torch.cuda.synchronize()
    for s, b, x in items:
        s.gmres_dev_async(b.data_ptr(), x.data_ptr())
    done = [s.wait() for s, b, x in items]
    for (s, b, x), (xa, ia), st in zip(items, alone, done):
        assert st.iters == ia and st.converged == 1
        assert torch.equal(x, xa)
    for s, _, _ in items:
        assert s.poll()
        s.close()


def test_async_busy_context_rejected():
    import torch

    from lowmachfhd_amd import solver as S

    (s, b, x), = _setup([Grid((64, 64, 64), 1 / 64)], S.FP64, S.GMRES_LEFT)
    s.gmres_dev_async(b.data_ptr(), x.data_ptr())
    with pytest.raises(S.LmfhdError) as ei:
        s.gmres_dev_async(b.data_ptr(), x.data_ptr())
    assert ei.value.code == -6
    st = s.wait()
    assert st.converged == 1
    s.close()
    torch.cuda.synchronize()
