"""Inhomogeneous walls on the GPU: apply_matrix with vel_bc_n / vel_bc_t
(apply_matrix.f90:24-40, 83-105) and convert_to_homogeneous (convert_to_homogeneous.f90:19)
against the oracle on random inputs (1e-13 relative; no-slip, slip and mixed walls, 2-D and 3-D,
every visc_type), the device-vector forms against the fab forms, and the reference's stored
Dirichlet accuracy goldens (Accuracy_tests/{2d,3d}_dis_accuracy_dirichlet.dat) at every stored
size: the discretization errors of the inhomogeneous apply_matrix on the exact solution, and
the global errors of the homogeneous GMRES solve of (tau_u, 0) and, for the Taylor vortex, of
convert_to_homogeneous followed by the solve; the global errors again through the mixed form."""
import json
import os

import numpy as np
import pytest
from fixtures import accuracy_problem, exact_wall_values
from test_gpu_parity import rel

from lowmachfhd_amd.grid import NO_SLIP_WALL, SLIP_WALL, Grid

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "dirichlet_goldens.json")))
N, S, P = NO_SLIP_WALL, SLIP_WALL, -1
CASES = [((12, 10, 8), (N, S, P), (S, N, P)), ((16, 12, 10), (N, N, N), (N, N, N)),
         ((12, 10, 8), (S, S, S), (S, S, S)), ((24, 20), (N, S), (S, N)), ((24, 20), (P, N), (P, N))]


def _solver(g, vt, **kn):
    from lowmachfhd_amd import solver as Sv

    return Sv.StokesSolver(g, visc_type=vt, **kn)


@pytest.mark.parametrize("vt", [1, -1, 2, -2, 3, -3])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_inhomogeneous_vs_oracle(case, vt):
    import torch
    from oracle import Oracle

    dims, lo, hi = CASES[case]
    g = Grid(dims, 1 / dims[0], bc_lo=lo, bc_hi=hi)
    rng = np.random.default_rng(abs(vt) + 3 + case)
    o = Oracle(g, visc_type=vt)
    s = _solver(g, vt)
    beta = 1 + 0.5 * rng.random(g.cc_shape()) if vt < 0 else np.full(g.cc_shape(), 1.3)
    bed = [1 + 0.5 * rng.random(g.edge_shape(e)) if vt < 0 else np.full(g.edge_shape(e), 1.3)
           for e in range(g.nedge)]
    gam = 0.3 * rng.random(g.cc_shape()) if vt < 0 else np.full(g.cc_shape(), 0.4)
    args = ([1 + rng.random(g.face_shape(a)) for a in range(g.dm)], beta, bed, gam, 2.0)
    o.set_coeffs(*args)
    s.set_coeffs(*args)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    p = rng.standard_normal(g.cc_shape())
    vbn = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    vbt = [rng.standard_normal(s.vel_bc_t_shape(q)) for q in range(len(s.vel_bc_t_pairs()))]
    b1, b2 = o.apply_matrix_bc(u, p, vbn, vbt), s.apply_matrix_bc(u, p, vbn, vbt)
    assert rel(b2[0], b1[0]) < 1e-13 and rel([b2[1]], [b1[1]]) < 1e-13
    c1, c2 = o.convert_to_homogeneous(u, p, vbn, vbt), s.convert_to_homogeneous(u, p, vbn, vbt)
    assert rel(c2[0], c1[0]) < 1e-13 and rel([c2[1]], [c1[1]]) < 1e-13
    # device-vector forms
    dev = lambda h: torch.from_numpy(np.ascontiguousarray(h)).to("cuda:0")  # noqa: E731
    x = dev(s.pack(u, p))
    vn = dev(s.pack(vbn, np.zeros(g.cc_shape())))
    vt_ = dev(np.concatenate([v.ravel() for v in vbt]))
    assert vt_.numel() == s.vel_bc_t_len
    b = torch.zeros_like(x)
    torch.cuda.synchronize()
    s.apply_matrix_bc_dev(x.data_ptr(), b.data_ptr(), vn.data_ptr(), vt_.data_ptr())
    s.sync()
    bu, bp = s.unpack(b.cpu().numpy())
    assert rel(bu, b2[0]) < 1e-14 and rel([bp], [b2[1]]) < 1e-14
    h = x.clone()
    s.convert_to_homogeneous_dev(h.data_ptr(), vn.data_ptr(), vt_.data_ptr())
    s.sync()
    hu, hp = s.unpack(h.cpu().numpy())
    assert rel(hu, c2[0]) < 1e-14 and rel([hp], [c2[1]]) < 1e-14
    s.close()


def _zero_walls(g, u):
    for a in range(g.dm):
        v = np.moveaxis(u[a], 2 - a, 0)
        v[0] = 0.0
        v[-1] = 0.0
    return u


@pytest.mark.parametrize("r", GOLD, ids=lambda r: f"{r['dim']}d-n{r['n']}-vt{r['visc_type']}-sol{r['prob_sol']}")
def test_dirichlet_goldens(r):
    dm, n, lo = r["dim"], r["n"], r["prob_lo"]
    g = Grid((n,) * dm, 1.0 / n, bc_lo=(N,) * dm, bc_hi=(N,) * dm, prob_lo=(lo, lo, lo))
    Pb = accuracy_problem(g, r["visc_type"], 1, r["prob_sol"])
    s = _solver(g, r["visc_type"], gmres_rel_tol=1e-12, gmres_max_inner=50, gmres_max_outer=20, gmres_max_iter=400)
    s.set_coeffs(Pb["alpha"], Pb["beta"], Pb["beta_ed"], Pb["gamma"], 0.0)
    vbn, vbt = exact_wall_values(g, r["prob_sol"])
    bu, bp = s.apply_matrix_bc(Pb["u"], Pb["p"], vbn, vbt)
    rhs = _zero_walls(g, [f.copy() for f in Pb["f"]])
    tu = [rhs[a] - bu[a] for a in range(dm)]
    tp = Pb["fp"] - bp
    for a, c in enumerate("UVW"[:dm] + "P"):
        got = float(np.abs(tu[a] if c != "P" else tp).max())
        gold = r[c]
        print(f"{c}: {got:.15e} (stored {gold:.15e})", flush=True)
        assert (got < 1e-9) if gold < 1e-9 else abs(got - gold) <= 1e-8 * gold
    eu, ep, st = s.gmres(tu, 0.0 * tp)
    assert st.converged == 1
    for a, c in enumerate("UVW"[:dm]):
        got, gold = float(np.abs(eu[a]).max()), r["g" + c]
        print(f"global {c}: {got:.15e} (stored {gold:.15e})", flush=True)
        assert abs(got - gold) <= 1e-6 * gold
    if r["prob_sol"] == 1:
        hu, hp = s.convert_to_homogeneous(rhs, Pb["fp"], vbn, vbt)
        xu, xp, st2 = s.gmres(hu, hp)
        assert st2.converged == 1
        err = _zero_walls(g, [xu[a] - Pb["u"][a] for a in range(dm)])
        for a, c in enumerate("UVW"[:dm]):
            got, gold = float(np.abs(err[a]).max()), r["g" + c]
            assert abs(got - gold) <= 1e-6 * gold, ("convert " + c, got, gold)
    s.close()


@pytest.mark.parametrize("r", [r for r in GOLD if r["n"] == 32],
                         ids=lambda r: f"{r['dim']}d-vt{r['visc_type']}-sol{r['prob_sol']}")
def test_dirichlet_global_error_mixed(r):
    """The stored global errors through the mixed form (FP32 multigrid inside FP64 FGMRES, true
    relative residual 1e-12): the same discrete solution, so the same error to 1e-6."""
    from lowmachfhd_amd import solver as Sv

    dm, n, lo = r["dim"], r["n"], r["prob_lo"]
    g = Grid((n,) * dm, 1.0 / n, bc_lo=(N,) * dm, bc_hi=(N,) * dm, prob_lo=(lo, lo, lo))
    Pb = accuracy_problem(g, r["visc_type"], 1, r["prob_sol"])
    s = _solver(g, r["visc_type"], precision=Sv.MIXED, krylov=Sv.FGMRES, gmres_rel_tol=1e-12, gmres_max_inner=50,
                gmres_max_outer=20, gmres_max_iter=400)
    s.set_coeffs(Pb["alpha"], Pb["beta"], Pb["beta_ed"], Pb["gamma"], 0.0)
    vbn, vbt = exact_wall_values(g, r["prob_sol"])
    bu, bp = s.apply_matrix_bc(Pb["u"], Pb["p"], vbn, vbt)
    rhs = _zero_walls(g, [f.copy() for f in Pb["f"]])
    tu = [rhs[a] - bu[a] for a in range(dm)]
    eu, ep, st = s.gmres(tu, 0.0 * Pb["fp"])
    assert st.converged == 1 and st.true_rel_resid <= 1e-12
    for a, c in enumerate("UVW"[:dm]):
        got, gold = float(np.abs(eu[a]).max()), r["g" + c]
        print(f"global {c}: {got:.15e} (stored {gold:.15e})", flush=True)
        assert abs(got - gold) <= 1e-6 * gold
    s.close()
