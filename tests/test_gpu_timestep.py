"""The momentum time steps on the GPU (lmfhd_advance_inertial_dev / _overdamped_dev,
lmfhd_advective_m_fluxdiv) against the oracle's restatement (oracle/timestep.py), on the same
inputs: variable density and viscosity (visc_type -1), a stochastic forcing, periodic and walled
boxes in 2-D and 3-D.  Bar: iteration counts within +-1 per solve, u and p within 1e-10 relative
after one and after three steps; the advective term alone within 1e-14.  At 64^3 a
size-independent property: a shear mode decays by the exact Crank-Nicolson factor."""
import numpy as np
import pytest

from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid

pytestmark = pytest.mark.gpu

TP = 2 * np.pi
GRIDS = {
    "per3": lambda: Grid((24, 24, 24), 1 / 24),
    "zwall3": lambda: Grid((24, 20, 16), 1 / 24, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL),
                           bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL)),
    "per2": lambda: Grid((48, 40), 1 / 48),
    "xwall2": lambda: Grid((48, 40), 1 / 48, bc_lo=(NO_SLIP_WALL, PERIODIC), bc_hi=(NO_SLIP_WALL, PERIODIC)),
    # no-slip x-walls and slip y-walls: both ghost rules in the advective and diffusive terms
    "mixwall3": lambda: Grid((20, 16, 12), 1 / 20, bc_lo=(NO_SLIP_WALL, SLIP_WALL, PERIODIC),
                             bc_hi=(NO_SLIP_WALL, SLIP_WALL, PERIODIC)),
}


def rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


def _problem(g, seed):
    rng = np.random.default_rng(seed)
    dm = g.dm
    u = [rng.standard_normal(g.face_shape(a)) for a in range(dm)]
    rho = [1 + 0.5 * rng.random(g.face_shape(a)) for a in range(dm)]
    eta = 1 + 0.5 * rng.random(g.cc_shape())
    eta_ed = [1 + 0.5 * rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    kap = np.zeros(g.cc_shape())
    p = rng.standard_normal(g.cc_shape())
    f1 = [0.3 * rng.standard_normal(g.face_shape(a)) for a in range(dm)]
    f2 = [0.3 * rng.standard_normal(g.face_shape(a)) for a in range(dm)]
    return u, rho, eta, eta_ed, kap, p, f1, f2


def _pair(g, **extra):
    from oracle import Oracle

    from lowmachfhd_amd import solver as S

    kn = dict(gmres_max_inner=30, gmres_rel_tol=1e-11, **extra)
    return (Oracle(g, visc_type=-1, **kn),
            S.StokesSolver(g, visc_type=-1, precision=S.FP64, krylov=S.GMRES_LEFT, **kn))


@pytest.mark.parametrize("name", list(GRIDS))
def test_advective_vs_oracle(name):
    from oracle import Oracle

    from lowmachfhd_amd import solver as S

    g = GRIDS[name]()
    u, rho, *_ = _problem(g, 1)
    s = S.StokesSolver(g, visc_type=-1)
    A2 = s.advective_m_fluxdiv(u, rho)
    A3 = s.advective_m_fluxdiv(u, rho, out=A2)  # increment: A2 - sum = 2 A2 (wall faces 0)
    s.close()
    A1 = Oracle(g, visc_type=-1).advective_m_fluxdiv(u, rho)
    assert rel(A2, A1) < 1e-14
    assert rel(A3, [2 * x for x in A1]) < 1e-14


OPTIONS = {"default": {}, "visc_schur": dict(visc_schur_approx=1), "lower_tri": dict(precon_type=2),
           "upper_tri_neg": dict(precon_type=-3)}


@pytest.mark.parametrize("name", list(GRIDS))
@pytest.mark.parametrize("kind", ["inertial", "overdamped"])
def test_step_vs_oracle(name, kind, opt="default"):
    import torch

    from oracle import timestep as TS

    g = GRIDS[name]()
    u, rho, eta, eta_ed, kap, p, f1, f2 = _problem(g, 2)
    o, s = _pair(g, **OPTIONS[opt])
    dt = 0.05
    if kind == "inertial":
        s.set_coeffs(rho, 0.5 * eta, [0.5 * e for e in eta_ed], 0.5 * kap, 1.0 / dt)
    else:
        s.set_coeffs(rho, eta, eta_ed, kap, 0.0)
    dev = lambda v: torch.from_numpy(np.ascontiguousarray(v)).to("cuda:0")  # noqa: E731
    x = dev(s.pack(u, p))
    F1, F2 = dev(s.pack(f1)), dev(s.pack(f2))
    uo, po = u, p
    for step in range(3):
        if kind == "inertial":
            uo, po, a1, a2 = TS.advance_inertial(o, g, dt, rho, eta, eta_ed, kap, uo, po, f1, f2, gmres_rel_tol=1e-11)
            b1, b2 = s.advance_inertial_dev(dt, x.data_ptr(), F1.data_ptr(), F2.data_ptr())
        else:
            uo, po, a1, a2 = TS.advance_overdamped(o, g, rho, eta, eta_ed, kap, uo, po, f1, f2, gmres_rel_tol=1e-11)
            b1, b2 = s.advance_overdamped_dev(x.data_ptr(), F1.data_ptr(), F2.data_ptr())
        ug, pg = s.unpack(x.cpu().numpy())
        eu, ep = rel(ug, uo), rel([pg - pg.mean()], [po - po.mean()])
        print(f"{name} {kind} step {step}: iterations {a1.total_iter}/{b1.iters} {a2.total_iter}/{b2.iters} "
              f"u {eu:.1e} p {ep:.1e}", flush=True)
        assert b1.converged == 1 and b2.converged == 1
        assert abs(a1.total_iter - b1.iters) <= 1 and abs(a2.total_iter - b2.iters) <= 1
        assert eu < 1e-10 and ep < 1e-10
    s.close()


@pytest.mark.parametrize("opt", ["visc_schur", "lower_tri", "upper_tri_neg"])
@pytest.mark.parametrize("name", ["per3", "zwall3"])
def test_step_options_vs_oracle(name, opt):
    """The inertial step with other preconditioners (the step's solves use the context's)."""
    test_step_vs_oracle(name, "inertial", opt)


def test_inertial_shear_mode_64():
    import torch

    from lowmachfhd_amd import solver as S

    n, k, A, eta, dt = 64, 3, 0.7, 0.2, 0.002
    g = Grid((n, n, n), 1 / n)
    s = S.StokesSolver(g, visc_type=1, precision=S.MIXED, krylov=S.FGMRES, gmres_max_inner=30, gmres_rel_tol=1e-12)
    rho = [np.ones(g.face_shape(a)) for a in range(3)]
    s.set_coeffs(rho, 0.5 * eta * np.ones(g.cc_shape()), [0.5 * eta * np.ones(g.edge_shape(e)) for e in range(3)],
                 np.zeros(g.cc_shape()), 1.0 / dt)
    Y = g.face_coords(0)[1]
    u = [A * np.sin(TP * k * Y), np.zeros(g.face_shape(1)), np.zeros(g.face_shape(2))]
    x = torch.from_numpy(s.pack(u, np.zeros(g.cc_shape()))).to("cuda:0")
    mu = 4 * n * n * np.sin(np.pi * k / n) ** 2
    fac = (1 / dt - 0.5 * eta * mu) / (1 / dt + 0.5 * eta * mu)
    for step in range(1, 4):
        s.advance_inertial_dev(dt, x.data_ptr())
        ug, pg = s.unpack(x.cpu().numpy())
        err = np.abs(ug[0] - fac ** step * u[0]).max() / A
        print(f"step {step}: factor {fac:.6f}, error {err:.1e}", flush=True)
        assert err < 1e-9 and max(np.abs(v).max() for v in ug[1:]) < 1e-9 * A
    s.close()


def test_step_argument_errors():
    import torch

    from lowmachfhd_amd import solver as S

    g = Grid((16, 16, 16), 1 / 16)
    s = S.StokesSolver(g, visc_type=1)
    one = np.ones(g.cc_shape())
    s.set_coeffs([np.ones(g.face_shape(a)) for a in range(3)], one, [np.ones(g.edge_shape(e)) for e in range(3)],
                 0 * one, 10.0)
    x = torch.zeros(int(s.vec_len), dtype=torch.float64, device="cuda:0")
    with pytest.raises(S.LmfhdError):
        s.advance_inertial_dev(0.05, x.data_ptr())     # theta is 10, not 1/0.05
    with pytest.raises(S.LmfhdError):
        s.advance_overdamped_dev(x.data_ptr())         # theta must be 0
    s.advance_inertial_dev(0.1, x.data_ptr())          # theta = 1/dt: a zero state stays zero
    assert float(x.abs().max()) == 0.0
    s.close()
