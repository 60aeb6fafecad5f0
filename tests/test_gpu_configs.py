"""The benchmark configurations at their full sizes, on the GPU.

  * config 0 (2-D 64^2 periodic, rho = eta = 1, dt = 0.1, stochastic RHS) and config 1 (3-D
    128^3 periodic, variable viscosity, stochastic + Fourier RHS): the GPU FP64 solve against
    the FP64 oracle on identical inputs -- iteration count within +-1, u and p within 1e-10
    relative; the mixed form (FP32 multigrid in FP64 FGMRES) at the 1e-9 true residual with
    its error against the oracle stated (bound 1e-7).
  * config 2 (3-D 256^3, no-slip z walls, variable density and viscosity): too large for the
    oracle in a test, so the mixed form is checked against the GPU FP64 form, whose parity
    with the oracle the other tests pin: true residual <= 1e-9 and the two solutions within
    1e-7 of each other, and the discrete divergence of u at the solver tolerance.
"""
import numpy as np
import pytest

from lowmachfhd_amd import problems
from lowmachfhd_amd.grid import Grid

pytestmark = pytest.mark.gpu


def rel(a, b):
    if isinstance(a, (list, tuple)):
        return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _solver(g, vt, mixed=False, **kn):
    from lowmachfhd_amd import solver as S

    prec, kry = (S.MIXED, S.FGMRES) if mixed else (S.FP64, S.GMRES_LEFT)
    kn.setdefault("gmres_max_inner", 30)
    return S.StokesSolver(g, visc_type=vt, precision=prec, krylov=kry, **kn)


def _check_vs_oracle(prob, label):
    import oracle

    g, vt = prob["grid"], prob["visc_type"]
    args = (prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
    o = oracle.Oracle(g, visc_type=vt, gmres_max_inner=30, omp_threads=16)
    o.set_coeffs(*args)
    xu1, xp1, st1, rc = o.gmres(prob["b_u"], prob["b_p"])
    assert rc == 0 and st1.converged == 1
    s = _solver(g, vt)
    s.set_coeffs(*args)
    xu2, xp2, st2 = s.gmres(prob["b_u"], prob["b_p"])
    print(f"{label}: oracle {st1.total_iter} iterations, GPU {st2.iters}; "
          f"u {rel(xu2, xu1):.2e} p {rel(xp2, xp1):.2e}", flush=True)
    assert st2.converged == 1 and abs(st2.iters - st1.total_iter) <= 1
    assert rel(xu2, xu1) < 1e-10 and rel(xp2, xp1) < 1e-10
    m = _solver(g, vt, mixed=True)
    m.set_coeffs(*args)
    xu3, xp3, st3 = m.gmres(prob["b_u"], prob["b_p"])
    eu, ep = rel(xu3, xu1), rel(xp3 - xp3.mean(), xp1 - xp1.mean())
    print(f"{label} mixed: {st3.iters} iterations, true residual {st3.true_rel_resid:.2e}, "
          f"error vs oracle u {eu:.2e} p {ep:.2e}", flush=True)
    assert st3.converged == 1 and st3.true_rel_resid <= 1e-9
    assert eu < 1e-7 and ep < 1e-7
    s.close()
    m.close()


def test_config0_vs_oracle():
    g = Grid((64, 64), 1.0 / 64)
    h = _solver(g, 1)
    prob = problems.config0(h, h)
    h.close()
    _check_vs_oracle(prob, "config 0")


def test_config1_vs_oracle():
    g = Grid((128, 128, 128), 1.0 / 128)
    h = _solver(g, -1, gmres_max_inner=1)
    prob = problems.config1_like(g, seed=11, averager=h)
    h.close()
    _check_vs_oracle(prob, "config 1")


def test_config2_mixed_vs_fp64():
    g = problems.config4_grid(256, True)
    h = _solver(g, -1, gmres_max_inner=1)
    prob = problems.config2_like(g, seed=12, averager=h)
    h.close()
    args = (prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
    s = _solver(g, -1)
    s.set_coeffs(*args)
    xu1, xp1, st1 = s.gmres(prob["b_u"], prob["b_p"])
    s.close()
    m = _solver(g, -1, mixed=True)
    m.set_coeffs(*args)
    xu2, xp2, st2 = m.gmres(prob["b_u"], prob["b_p"])
    div = m.compute_div(xu2)
    m.close()
    eu, ep = rel(xu2, xu1), rel(xp2 - xp2.mean(), xp1 - xp1.mean())
    print(f"config 2: FP64 {st1.iters} iterations, mixed {st2.iters}; mixed true residual "
          f"{st2.true_rel_resid:.2e}; mixed vs FP64 u {eu:.2e} p {ep:.2e}; max|div u| {np.abs(div).max():.2e}",
          flush=True)
    assert st1.converged == 1 and st2.converged == 1 and st2.true_rel_resid <= 1e-9
    assert eu < 1e-7 and ep < 1e-7
    # -div u = b_p = 0 up to the solver tolerance relative to the scale of D applied to u
    scale = max(np.abs(x).max() for x in xu2) / g.dx
    assert np.abs(div).max() <= 1e-7 * scale
