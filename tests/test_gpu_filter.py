"""filtering_width 1, 2, 4 on the GPU (filter_stoch_m_flux, stochastic_m_fluxdiv.f90:604-778,
applied to the scaled stresses of stochastic_m_fluxdiv at :172-175 / :226-231): the fab entry
point against the oracle on the same noise (1e-13 relative), on periodic 2-D and 3-D boxes
including one whose shortest side equals the width; the on-device noise path against the fab
path fed the same normals, in both stress forms and with `increment`."""
import numpy as np
import pytest
from test_gpu_parity import _philox_normals, rel

from lowmachfhd_amd.grid import Grid

pytestmark = pytest.mark.gpu

DIMS = [(16, 12, 10), (24, 20), (8, 8, 4), (32, 32, 32)]


def _solver(g, w):
    from lowmachfhd_amd import solver as S

    return S.StokesSolver(g, visc_type=-1, filtering_width=w)


@pytest.mark.parametrize("w", [1, 2, 4])
@pytest.mark.parametrize("dims", DIMS)
def test_filtered_fluxdiv_vs_oracle(dims, w):
    from oracle import Oracle

    g = Grid(dims, 1 / dims[0])
    rng = np.random.default_rng(w + 7)
    o = Oracle(g, visc_type=-1, filtering_width=w)
    s = _solver(g, w)
    Wcc = [rng.standard_normal(g.cc_shape()) for _ in range(g.dm)]
    Wed = [rng.standard_normal(g.edge_shape(q // 2)) for q in range(2 * g.nedge)]
    eta = 1 + rng.random(g.cc_shape())
    eed = [1 + rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    T = 1 + rng.random(g.cc_shape())
    Ted = [1 + rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    a = s.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.1, T_cc=T, T_ed=Ted)
    b = o.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.1, T_cc=T, T_ed=Ted)
    print(f"{dims} width {w}: {rel(a, b):.1e}", flush=True)
    assert rel(a, b) < 1e-13
    # filtering changes the result (the unfiltered divergence differs by O(1))
    o0 = Oracle(g, visc_type=-1)
    assert rel(a, o0.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.1, T_cc=T, T_ed=Ted)) > 0.1
    s.close()


@pytest.mark.parametrize("form", [0, 1])
@pytest.mark.parametrize("w", [2, 4])
@pytest.mark.parametrize("dims", [(16, 12, 10), (24, 20)])
def test_filtered_device_noise(dims, w, form):
    import torch

    g = Grid(dims, 1 / dims[0])
    rng = np.random.default_rng(5)
    s = _solver(g, w)
    eta = 1 + 0.5 * rng.random(g.cc_shape())
    eta_ed = [1 + 0.5 * rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    one = [np.ones(g.face_shape(a)) for a in range(g.dm)]
    s.set_coeffs(one, eta, eta_ed, np.zeros(g.cc_shape()), 1.0)
    seed, stage = 12345, 3
    W_cc = [(np.sqrt(2.0) if form else 1.0) *
            _philox_normals(seed, stage, 0, g.ncells, a).reshape(g.cc_shape()) for a in range(g.dm)]
    W_ed = []
    for e in range(g.nedge):
        m = int(np.prod(g.edge_shape(e)))
        for q in range(2):
            W_ed.append(_philox_normals(seed, stage, 1 + e, m, q if form == 0 else 0).reshape(g.edge_shape(e)))
    ref = s.stochastic_m_fluxdiv(W_cc, W_ed, eta, eta_ed, 0.1, 1.0, 1.0)
    f = torch.full((int(s.vec_len),), 7.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    s.stochastic_m_fluxdiv_dev(seed, stage, f.data_ptr(), 0.1, 1.0, 1.0, stoch_stress_form=form)
    s.sync()
    fu, fp = s.unpack(f.cpu().numpy())
    assert rel(fu, ref) < 1e-12 and np.all(fp == 0.0)
    base = torch.from_numpy(rng.standard_normal(int(s.vec_len))).to("cuda:0")
    g_inc = base.clone()
    torch.cuda.synchronize()
    s.stochastic_m_fluxdiv_dev(seed, stage, g_inc.data_ptr(), 0.1, 1.0, 1.0, stoch_stress_form=form,
                               increment=True)
    s.sync()
    assert torch.allclose(g_inc, base + f, rtol=0, atol=1e-12 * float(f.abs().max()))
    s.close()
