"""The entry point that allocates nothing (lmfhd_stochastic_m_fluxdiv_dev_noalloc) against the
allocating one (lmfhd_stochastic_m_fluxdiv_dev), through the C ABI on the GPU: bit-identical
results on two calls in a row, the second on a smaller grid, with filtering on (the case that
stages through the library's fixed device array) and off; no device memory added to the context
by the no-allocation call; a grid too large for the staging array is refused before any launch."""
import numpy as np
import pytest

from lowmachfhd_amd.grid import Grid

pytestmark = pytest.mark.gpu


def _setup(dims, w):
    import torch

    from lowmachfhd_amd import solver as S

    g = Grid(dims, 1.0 / dims[0])
    s = S.StokesSolver(g, visc_type=-1, filtering_width=w)
    rng = np.random.default_rng(sum(dims) + w)
    eta = 1 + 0.5 * rng.random(g.cc_shape())
    eta_ed = [1 + 0.5 * rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    one = [np.ones(g.face_shape(a)) for a in range(g.dm)]
    s.set_coeffs(one, eta, eta_ed, np.zeros(g.cc_shape()), 1.0)
    n = s.vec_len
    return s, torch.zeros(n, dtype=torch.float64, device="cuda:0"), rng


@pytest.mark.parametrize("w", [2, 0])
def test_noalloc_matches_allocating_entry_point(w):
    import torch

    for dims in ((40, 36, 32), (16, 12, 10), (24, 20)):  # the second call on a smaller grid
        s, f0, rng = _setup(dims, w)
        base = torch.from_numpy(rng.standard_normal(f0.numel())).to("cuda:0")
        f_na = base.clone()
        f_al = base.clone()
        torch.cuda.synchronize()
        for stage, form, inc in ((3, 0, False), (4, 1, True)):
            before = s.device_bytes
            s.stochastic_m_fluxdiv_dev_noalloc(77, stage, f_na.data_ptr(), 0.05, k_B=1.5, stoch_stress_form=form,
                                               increment=inc)
            s.sync()
            assert s.device_bytes == before, "the no-allocation entry point grew the context"
            s.stochastic_m_fluxdiv_dev(77, stage, f_al.data_ptr(), 0.05, k_B=1.5, stoch_stress_form=form,
                                       increment=inc)
            s.sync()
            print(f"{dims} width {w} stage {stage}: equal {torch.equal(f_na, f_al)}", flush=True)
            assert torch.equal(f_na, f_al), (dims, w, stage)
            assert torch.isfinite(f_na).all()
        s.close()


def test_noalloc_refuses_grid_beyond_staging():
    from lowmachfhd_amd import solver as S

    s, f0, _ = _setup((96, 96, 96), 1)
    with pytest.raises(S.LmfhdError) as ei:
        s.stochastic_m_fluxdiv_dev_noalloc(1, 0, f0.data_ptr(), 0.05)
    assert ei.value.code == -1  # LMFHD_E_ARG
    s.close()
