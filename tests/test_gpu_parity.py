"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle on identical inputs.

Tolerances (relative, max-norm over all entries):
  * operators (stag_applyop, apply_matrix, grad, div, averaging, stochastic RHS): 1e-13;
  * multigrid V-cycles and the preconditioner in FP64: 1e-12;
  * FP64 GMRES solution u, p: 1e-10, GMRES iteration count within +-1;
  * mixed precision (FP32 multigrid inside FP64 Krylov): true residual <= 1e-9 relative and
    solution within 1e-7 of the FP64 oracle (the error is stated, not hidden).
"""
import json
import os

import numpy as np
import pytest

from fixtures import accuracy_problem
from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _solver_mod():
    from lowmachfhd_amd import solver

    return solver


def rel(a, b):
    if isinstance(a, (list, tuple)):
        num = max(np.abs(x - y).max() for x, y in zip(a, b))
        den = max(np.abs(y).max() for y in b)
    else:
        num, den = np.abs(a - b).max(), np.abs(b).max()
    return num / max(den, 1e-300)


GRIDS = {
    "per3": lambda: Grid((16, 16, 16), 1 / 16),
    "zwall3": lambda: Grid((16, 12, 16), 1 / 16, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL)),
    "mixwall3": lambda: Grid((8, 16, 12), 1 / 16, bc_lo=(NO_SLIP_WALL, PERIODIC, SLIP_WALL), bc_hi=(NO_SLIP_WALL, PERIODIC, SLIP_WALL)),
    "per2": lambda: Grid((32, 32), 1 / 32),
    "ywall2": lambda: Grid((32, 24), 1 / 32, bc_lo=(PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, NO_SLIP_WALL)),
    "odd3": lambda: Grid((12, 12, 12), 1 / 12),  # coarsest level 3^3: two-phase smoothing
    # larger grids: the fine levels run as separate launches, the coarse ones in one workgroup
    "per32": lambda: Grid((32, 32, 32), 1 / 32),
    "zwall24": lambda: Grid((24, 24, 24), 1 / 24, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL)),
    "per2_64x48": lambda: Grid((64, 48), 1 / 64),
    # walls in x (no-slip) and z (slip) on a level smoothed by the two-colour tile kernel
    "mixwall32": lambda: Grid((32, 32, 24), 1 / 32, bc_lo=(NO_SLIP_WALL, PERIODIC, SLIP_WALL), bc_hi=(NO_SLIP_WALL, PERIODIC, SLIP_WALL)),
}


def dupsync(g, x, nodal_dims):
    for d in nodal_dims:
        if g.bc_lo[d] == PERIODIC:
            ax = 2 - d
            i0 = [slice(None)] * 3
            iN = [slice(None)] * 3
            i0[ax], iN[ax] = 0, -1
            x[tuple(iN)] = x[tuple(i0)]
    return x


def consistent_u(g, u):
    """Equal duplicated periodic faces, zero wall-normal faces (vectors of the solution space)."""
    for a in range(g.dm):
        dupsync(g, u[a], [a])
        if g.bc_lo[a] != PERIODIC:
            ax = 2 - a
            i0 = [slice(None)] * 3
            iN = [slice(None)] * 3
            i0[ax], iN[ax] = 0, -1
            u[a][tuple(i0)] = 0
            u[a][tuple(iN)] = 0
    return u


def random_coeffs(g, rng, variable):
    beta = 1 + 0.5 * rng.random(g.cc_shape()) if variable else np.full(g.cc_shape(), 1.3)
    gamma = 0.3 * rng.random(g.cc_shape()) if variable else np.full(g.cc_shape(), 0.7)
    alpha = [dupsync(g, 1 + 0.5 * rng.random(g.face_shape(a)), [a]) for a in range(g.dm)]
    bed = [dupsync(g, 1 + 0.5 * rng.random(g.edge_shape(e)), g.edge_nodal(e)) for e in range(g.nedge)]
    return alpha, beta, bed, gamma


def pair(g, vt, theta, rng, precision=0, krylov=0, **kn):
    import oracle

    S = _solver_mod()
    al, be, bed, ga = random_coeffs(g, rng, vt < 0)
    okn = {}
    for k, v in kn.items():
        if k.startswith("stag_mg_"):
            okn[k.replace("stag_mg_", "stag_").replace("nsmooths_", "nu_")] = v
        elif k.startswith("mg_"):
            okn[k.replace("mg_", "mac_").replace("nsmooths_", "nu_")] = v
        else:
            okn[k] = v
    o = oracle.Oracle(g, visc_type=vt, **okn)
    s = S.StokesSolver(g, visc_type=vt, precision=precision, krylov=krylov, **kn)
    o.set_coeffs(al, be, bed, ga, theta)
    s.set_coeffs(al, be, bed, ga, theta)
    o.coeffs = (al, be, bed, ga, theta)
    return o, s


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("vt", [1, 2, 3, -1, -2, -3])
def test_stag_applyop_all_colors(gname, vt):
    g = GRIDS[gname]()
    rng = np.random.default_rng(abs(vt) * 7 + len(gname))
    o, s = pair(g, vt, 7.0, rng)
    phi = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    for color in range(0, 2 * g.dm + 1):
        L0 = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
        a = o.stag_applyop(phi, color, Lphi=[x.copy() for x in L0])
        b = s.stag_applyop(phi, color, Lphi=L0)
        assert rel(b, a) < 1e-13, (gname, vt, color)
    p = rng.standard_normal(g.cc_shape())
    bu1, bp1 = o.apply_matrix(phi, p)
    bu2, bp2 = s.apply_matrix(phi, p)
    assert rel(bu2, bu1) < 1e-13 and rel(bp2, bp1) < 1e-13


@pytest.mark.parametrize("gname", list(GRIDS))
def test_grad_div_average(gname):
    g = GRIDS[gname]()
    rng = np.random.default_rng(5)
    o, s = pair(g, -1, 1.0, rng)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    p = rng.standard_normal(g.cc_shape())
    assert rel(s.compute_div(u), o.compute_div(u)) < 1e-13
    assert rel(s.compute_grad(p), o.compute_grad(p)) < 1e-13
    nx, ny, nz = g.n3
    ccg = rng.random((nz + 2 if g.dm == 3 else 1, ny + 2, nx + 2))
    assert rel(s.average_cc_to_face(ccg), o.average_cc_to_face(ccg)) < 1e-14
    assert rel(s.average_cc_to_edge(ccg), o.average_cc_to_edge(ccg)) < 1e-14


@pytest.mark.parametrize("gname", list(GRIDS))
def test_stochastic_m_fluxdiv(gname):
    g = GRIDS[gname]()
    rng = np.random.default_rng(9)
    o, s = pair(g, -1, 1.0, rng)
    Wcc = [rng.standard_normal(g.cc_shape()) for _ in range(g.dm)]
    Wed = [rng.standard_normal(g.edge_shape(q // 2)) for q in range(2 * g.nedge)]
    eta = 1 + rng.random(g.cc_shape())
    eed = [1 + rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    T = 1 + rng.random(g.cc_shape())
    Ted = [1 + rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    assert rel(s.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.1, T_cc=T, T_ed=Ted),
               o.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.1, T_cc=T, T_ed=Ted)) < 1e-13
    assert rel(s.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.05, 1.0, 2.0),
               o.stochastic_m_fluxdiv(Wcc, Wed, eta, eed, 0.05, 1.0, 2.0)) < 1e-13


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("vt", [1, -1, -2, 3])
def test_stag_mg_solver(gname, vt):
    g = GRIDS[gname]()
    rng = np.random.default_rng(11 + vt)
    for nv in (1, 3):
        o, s = pair(g, vt, 5.0, rng, stag_mg_max_vcycles=nv, stag_mg_rel_tol=1e-30)
        rhs = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
        phi0 = consistent_u(g, [0.1 * rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
        assert rel(s.stag_mg_solver(rhs, phi0), o.stag_mg_solve(rhs, phi0)) < 1e-12, (gname, vt, nv)


@pytest.mark.parametrize("gname", list(GRIDS))
def test_mac_mg_and_precon(gname):
    g = GRIDS[gname]()
    rng = np.random.default_rng(21)
    o, s = pair(g, -1, 5.0, rng)
    r = rng.standard_normal(g.cc_shape())
    assert rel(s.mac_mg_solver(r), o.mac_mg_solve(r)) < 1e-12
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = rng.standard_normal(g.cc_shape())
    xu1, xp1 = o.apply_precon(bu, bp)
    xu2, xp2 = s.apply_precon(bu, bp)
    assert rel(xu2, xu1) < 1e-12 and rel(xp2, xp1) < 1e-12


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("vt", [1, -1, -2])
def test_gmres_fp64(gname, vt):
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
    assert st2.true_rel_resid < 1e-7


@pytest.mark.parametrize("gname", ["per3", "zwall3", "per2", "odd3", "per32", "mixwall32"])
def test_mixed_precision(gname):
    """FP32 multigrid inside FP64 FGMRES: true residual at the FP64 tolerance, error vs oracle."""
    S = _solver_mod()
    g = GRIDS[gname]()
    rng = np.random.default_rng(41)
    o, s = pair(g, -1, 5.0, rng, precision=S.MIXED, krylov=S.FGMRES, gmres_max_inner=30)
    pair_coeffs = [o.coeffs]
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = np.zeros(g.cc_shape())
    import oracle

    o = oracle.Oracle(g, visc_type=-1, gmres_max_inner=30, gmres_rel_tol=1e-12)
    o.set_coeffs(*pair_coeffs[0])
    xu1, xp1, st1, rc = o.gmres(bu, bp)
    xu2, xp2, st2 = s.gmres(bu, bp)
    assert st2.converged == 1 and st2.true_rel_resid <= 1e-9
    assert rel(xu2, xu1) < 1e-7
    assert rel(xp2 - xp2.mean(), xp1 - xp1.mean()) < 1e-7


def test_accuracy_goldens_on_gpu():
    """The reference's truncation-error goldens through the HIP apply_matrix (32^3 .. 128^3)."""
    S = _solver_mod()
    recs = json.load(open(os.path.join(HERE, "golden", "accuracy_goldens.json")))
    for r in recs:
        if r["dim"] == 3 and r["n"] == 128 and r["visc_type"] not in (1, -1):
            continue
        n, dim = r["n"], r["dim"]
        g = Grid((n,) * dim, 1.0 / n)
        pr = accuracy_problem(g, r["visc_type"], r["prob_coeff"])
        s = S.StokesSolver(g, visc_type=r["visc_type"])
        s.set_coeffs(pr["alpha"], pr["beta"], pr["beta_ed"], pr["gamma"], 0.0)
        bu, bp = s.apply_matrix(pr["u"], pr["p"])
        errs = [np.abs(pr["f"][d] - bu[d]).max() for d in range(dim)] + [np.abs(pr["fp"] - bp).max()]
        tol = 32 * np.finfo(float).eps * n * n
        for e, c in zip(errs, "UVW"[:dim] + "P"):
            assert abs(e - r[c]) <= tol, (r, c, e)
        s.close()


def test_error_paths():
    S = _solver_mod()
    g = Grid((8, 8, 8), 1 / 8)
    rng = np.random.default_rng(3)
    o, s = pair(g, 1, 5.0, rng)
    # zero right-hand side: x = 0 in zero iterations (gmres.f90:162-190)
    xu, xp, st = s.gmres(s.zeros_u(), s.zeros_p())
    assert st.converged == 1 and st.iters == 0 and all(np.all(x == 0) for x in xu)
    # NaN in the right-hand side is reported, not propagated
    bu = [np.full(g.face_shape(a), np.nan) for a in range(3)]
    with pytest.raises(S.LmfhdError) as ei:
        s.gmres(bu, s.zeros_p())
    assert ei.value.code == -4
    # wrong shape
    with pytest.raises(S.LmfhdError):
        s.compute_div([np.zeros((3, 3, 3))] * 3)
    # gmres_max_iter = 1: stops unconverged, warning-only like the reference
    o2, s2 = pair(g, 1, 5.0, rng, gmres_max_iter=1, gmres_max_inner=30)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(3)])
    xu, xp, st = s2.gmres(bu, s2.zeros_p())
    assert st.iters == 1 and st.converged == 0


@pytest.mark.parametrize("gname", ["per3", "zwall3", "per2"])
@pytest.mark.parametrize("form", [0, 1])
def test_stochastic_rhs_reference_stream(gname, form):
    """The RHS built from the reference's own noise stream (HydroGrid MT19937, fill order of
    fill_m_stochastic) by the GPU equals the oracle's."""
    from lowmachfhd_amd import problems

    g = GRIDS[gname]()
    rng = np.random.default_rng(3)
    o, s = pair(g, -1, 1.0, rng)
    eta = 1 + 0.5 * rng.random(g.cc_shape())
    one = np.ones(g.cc_shape())
    kw = dict(noise_source="hydrogrid", stoch_stress_form=form)
    a = problems.assemble(g, -1, one, eta, 10.0, 0.1, 1234, s, s, **kw)
    b = problems.assemble(g, -1, one, eta, 10.0, 0.1, 1234, s, o, **kw)
    assert rel(a["b_u"], b["b_u"]) < 1e-13


def _philox_normals(seed, stage, array, npos, form_word):
    """Host mirror of the on-device noise: numpy's Philox4x64-10 (which advances its counter
    before each block) started one below (0, array, stage, 0), key (seed, 0x243F6A8885A308D3);
    uniform ((w >> 11) + 1/2) 2^-53; exact normal quantile."""
    from scipy.special import ndtri

    ctr = (array << 64) + (stage << 128) - 1
    bg = np.random.Philox(counter=ctr % (1 << 256), key=seed + (0x243F6A8885A308D3 << 64))
    w = bg.random_raw(4 * npos).reshape(npos, 4)[:, form_word]
    u = ((w >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    return ndtri(u)


@pytest.mark.parametrize("gname", ["per3", "zwall3", "per2", "mixwall32"])
@pytest.mark.parametrize("form", [0, 1])
def test_stochastic_device_noise(gname, form):
    """On-device noise (Philox, regenerated where the flux divergence needs it) equals the
    array path fed with the same normals generated on the host."""
    import torch

    g = GRIDS[gname]()
    rng = np.random.default_rng(8)
    o, s = pair(g, -1, 1.0, rng)
    eta = 1 + 0.5 * rng.random(g.cc_shape())
    eta_ed = [1 + 0.5 * rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    one = [np.ones(g.face_shape(a)) for a in range(g.dm)]
    s.set_coeffs(one, eta, eta_ed, np.zeros(g.cc_shape()), 1.0)
    seed, stage = 987654321, 5
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
    assert rel(fu, ref) < 1e-12
    assert np.all(fp == 0.0)
    # another stage gives other noise
    s.stochastic_m_fluxdiv_dev(seed, stage + 1, f.data_ptr(), 0.1)
    s.sync()
    fu2, _ = s.unpack(f.cpu().numpy())
    assert rel(fu2, ref) > 0.1


def test_stochastic_device_noise_increment():
    import torch

    g = GRIDS["zwall3"]()
    rng = np.random.default_rng(2)
    o, s = pair(g, -1, 1.0, rng)
    base = torch.from_numpy(rng.standard_normal(int(s.vec_len))).to("cuda:0")
    f = torch.zeros_like(base)
    torch.cuda.synchronize()
    s.stochastic_m_fluxdiv_dev(11, 0, f.data_ptr(), 0.1)
    g_inc = base.clone()
    torch.cuda.synchronize()
    s.stochastic_m_fluxdiv_dev(11, 0, g_inc.data_ptr(), 0.1, increment=True)
    s.sync()
    assert torch.allclose(g_inc, base + f, rtol=0, atol=1e-12 * float(f.abs().max()))


@pytest.mark.parametrize("gname", ["per3", "zwall3", "per2", "per32", "mixwall32"])
@pytest.mark.parametrize("pt", [-1, 2, -2, 3, -3, 4, -4, 5, -5])
def test_precon_types(gname, pt):
    """apply_precon for precon_type -1 and +-2 .. +-5 (apply_precon.f90:101-426) against the
    oracle, both theta > 0 and theta = 0."""
    g = GRIDS[gname]()
    rng = np.random.default_rng(50 + pt)
    for vt, theta in ((-1, 5.0), (3, 0.0)):
        o, s = pair(g, vt, theta, rng, precon_type=pt)
        bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
        bp = rng.standard_normal(g.cc_shape())
        xu1, xp1 = o.apply_precon(bu, bp)
        xu2, xp2 = s.apply_precon(bu, bp)
        assert rel(xu2, xu1) < 1e-12 and rel(xp2, xp1) < 1e-12, (gname, pt, vt)


@pytest.mark.parametrize("pt", [-1, 2, 3, 4, 5])
def test_gmres_precon_types(pt):
    g = GRIDS["zwall24"]()
    rng = np.random.default_rng(60 + pt)
    o, s = pair(g, -1, 5.0, rng, precon_type=pt, gmres_max_inner=30, gmres_max_iter=200)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = np.zeros(g.cc_shape())
    xu1, xp1, st1, rc = o.gmres(bu, bp)
    xu2, xp2, st2 = s.gmres(bu, bp)
    assert rc == 0 and st1.converged == 1 and st2.converged == 1
    assert abs(st2.iters - st1.total_iter) <= 1
    assert rel(xu2, xu1) < 1e-9 and rel(xp2 - xp2.mean(), xp1 - xp1.mean()) < 1e-9


def _brute_norm(g, u, p, pw):
    """sqrt(sum_d sum w u_d^2 + (pw |p|)^2), w = 1/2 on the faces 0 and n of component d's normal
    direction (norm_inner_product.f90:199-312; gmres.f90:148-151 scales the pressure norm by
    p_norm_weight)."""
    s = 0.0
    for d in range(g.dm):
        w = np.ones(g.face_shape(d))
        i0 = [slice(None)] * 3
        iN = [slice(None)] * 3
        i0[2 - d], iN[2 - d] = 0, -1
        w[tuple(i0)] = 0.5
        w[tuple(iN)] = 0.5
        s += float((w * u[d] * u[d]).sum())
    return float(np.sqrt(s + (pw * np.sqrt(float((p * p).sum()))) ** 2))


@pytest.mark.parametrize("gname", ["per3", "zwall3", "mixwall3", "per2", "ywall2"])
@pytest.mark.parametrize("pw", [1.0, 2.5])
def test_krylov_norms_match_brute_force(gname, pw):
    """The device inner products (one reduction tree per component, 1/2 face weights) against a
    numpy sum: |b| and |M^-1 b| of gmres (gmres.f90:137-151)."""
    g = GRIDS[gname]()
    rng = np.random.default_rng(41)
    o, s = pair(g, -1, 5.0, rng, p_norm_weight=pw)
    bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    bp = rng.standard_normal(g.cc_shape())
    _, _, st = s.gmres(bu, bp)
    assert abs(st.norm_b - _brute_norm(g, bu, bp, pw)) <= 1e-13 * st.norm_b
    pu, pp = o.apply_precon(bu, bp)
    assert abs(st.norm_pre_b - _brute_norm(g, pu, pp, pw)) <= 1e-11 * st.norm_pre_b
