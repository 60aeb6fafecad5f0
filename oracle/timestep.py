"""Momentum time steps for one species, restated from the reference on the oracle's operators.

TEST INFRASTRUCTURE ONLY (the checker of lmfhd_advance_inertial_dev /
lmfhd_advance_overdamped_dev); the product package never imports it.

The reference routines advance a multispecies low-Mach fluid.  For one species whose density,
viscosities and temperature do not change in time, with no gravity and homogeneous velocity
boundaries, their mass-flux, scalar and charge parts vanish and what remains is the momentum
update restated here, term by term in the reference's order:

  advance_inertial   -- src_lowMach/advance_timestep_inertial.f90:64-915 (Step 4, predictor
                        Crank-Nicolson, :242-520; Step 7, corrector, :640-890)
  advance_overdamped -- src_lowMach/advance_timestep_overdamped.f90:50-785 (predictor Stokes
                        solve, :200-375; corrector, :478-680)

Fields are numpy arrays in the valid-region layout of lowmachfhd_amd.grid (lists of face arrays
for velocities, one cell array for the pressure).  f1, f2 are stochastic flux divergences as
the reference's stochastic_m_fluxdiv returns them (see the C-ABI's description).
"""
from __future__ import annotations

import numpy as np

from lowmachfhd_amd.grid import is_wall


def _zero_walls(grid, u):
    """multifab_physbc_domainvel with homogeneous data: wall-normal faces of u set to 0."""
    out = [x.copy() for x in u]
    for a in range(grid.dm):
        ax = 2 - a  # numpy axis of direction a (arrays are (z, y, x), nz = 1 in 2-D)
        if is_wall(grid.bc_lo[a]):
            idx = [slice(None)] * 3
            idx[ax] = 0
            out[a][tuple(idx)] = 0.0
        if is_wall(grid.bc_hi[a]):
            idx = [slice(None)] * 3
            idx[ax] = -1
            out[a][tuple(idx)] = 0.0
    return out


def _coef(o, rho_fc, eta_cc, eta_ed, kappa_cc, theta, scale):
    o.set_coeffs(rho_fc, scale * eta_cc, [scale * e for e in eta_ed], scale * kappa_cc, theta)


def advance_inertial(o, grid, dt, rho_fc, eta_cc, eta_ed, kappa_cc, u, p, f1=None, f2=None,
                     gmres_rel_tol=1e-9, gmres_abs_tol=0.0):
    """One step; returns (u, p, stats_predictor, stats_corrector)."""
    dm = grid.dm
    theta = 1.0 / dt
    u = _zero_walls(grid, u)
    # ---- Step 4 (predictor): rhs_v = mold/dt - G pi + A^n + 1/2 A0 v + div Sigma - rho v/dt
    #      + 1/2 A0 v, zero on walls; rhs_p = D v; eta and kappa halved for the solve
    _coef(o, rho_fc, eta_cc, eta_ed, kappa_cc, theta, 1.0)
    mold = [rho_fc[a] * u[a] for a in range(dm)]
    rhs = [(1.0 / dt) * mold[a] + 0.0 * mold[a] for a in range(dm)]
    gp = o.compute_grad(p)
    rhs = [rhs[a] - gp[a] for a in range(dm)]
    adv = o.advective_m_fluxdiv(u, rho_fc)
    rhs = [rhs[a] + adv[a] for a in range(dm)]
    diff = o.diffusive_m_fluxdiv(u)
    rhs = [rhs[a] + 0.5 * diff[a] for a in range(dm)]
    if f1 is not None:
        rhs = [rhs[a] + f1[a] for a in range(dm)]
    mtemp = [rho_fc[a] * u[a] for a in range(dm)]
    rhs = [rhs[a] + (-1.0 / dt) * mtemp[a] for a in range(dm)]
    mtemp = o.diffusive_m_fluxdiv(u)
    rhs = [rhs[a] + 0.5 * mtemp[a] for a in range(dm)]
    rhs = _zero_walls(grid, rhs)
    rhs_p = np.zeros(grid.cc_shape()) + o.compute_div(u)
    _coef(o, rho_fc, eta_cc, eta_ed, kappa_cc, theta, 0.5)
    o.set_gmres_abs_tol(gmres_abs_tol)
    du, dp, st1, rc = o.gmres(rhs, rhs_p)
    assert rc == 0
    abs2 = max(gmres_abs_tol, st1.norm_pre_b * gmres_rel_tol)
    u = [u[a] + du[a] for a in range(dm)]
    p = p + dp
    # ---- Step 7 (corrector)
    _coef(o, rho_fc, eta_cc, eta_ed, kappa_cc, theta, 1.0)
    rhs = [(1.0 / dt) * mold[a] + 0.0 * mold[a] for a in range(dm)]
    gp = o.compute_grad(p)
    rhs = [rhs[a] - gp[a] for a in range(dm)]
    adv = o.advective_m_fluxdiv(u, rho_fc, out=adv)
    rhs = [rhs[a] + 0.5 * adv[a] for a in range(dm)]
    rhs = [rhs[a] + 0.5 * diff[a] for a in range(dm)]
    if f1 is not None:
        stoch = [f1[a] + (f2[a] if f2 is not None else f1[a]) for a in range(dm)]
        rhs = [rhs[a] + 0.5 * stoch[a] for a in range(dm)]
    mtemp = [rho_fc[a] * u[a] for a in range(dm)]
    rhs = [rhs[a] + (-1.0 / dt) * mtemp[a] for a in range(dm)]
    diff = o.diffusive_m_fluxdiv(u)
    rhs = [rhs[a] + 0.5 * diff[a] for a in range(dm)]
    rhs = _zero_walls(grid, rhs)
    rhs_p = np.zeros(grid.cc_shape()) + o.compute_div(u)
    _coef(o, rho_fc, eta_cc, eta_ed, kappa_cc, theta, 0.5)
    o.set_gmres_abs_tol(abs2)
    du, dp, st2, rc = o.gmres(rhs, rhs_p)
    assert rc == 0
    o.set_gmres_abs_tol(gmres_abs_tol)
    u = [u[a] + du[a] for a in range(dm)]
    p = p + dp
    return u, p, st1, st2


def advance_overdamped(o, grid, rho_fc, eta_cc, eta_ed, kappa_cc, u, p, f1=None, f2=None,
                       gmres_rel_tol=1e-9, gmres_abs_tol=0.0):
    """One step (theta = 0); returns (u, p, stats_predictor, stats_corrector)."""
    dm = grid.dm
    o.set_coeffs(rho_fc, eta_cc, eta_ed, kappa_cc, 0.0)
    u = _zero_walls(grid, u)
    abs_tol = gmres_abs_tol
    stats = []
    for stage, f in enumerate((f1, f2)):
        gp = o.compute_grad(p)
        rhs = [-1.0 * gp[a] + 0.0 * gp[a] for a in range(dm)]      # saxpy_5(rhs, -1, gp, 0, gp)
        if f is not None:
            rhs = [rhs[a] + f[a] for a in range(dm)]                # stochastic, increment
        rhs = o.diffusive_m_fluxdiv(u, out=rhs)                     # + A0 v, increment
        rhs_p = np.zeros(grid.cc_shape()) + o.compute_div(u)
        o.set_gmres_abs_tol(abs_tol)
        du, dp, st, rc = o.gmres(rhs, rhs_p)
        assert rc == 0
        stats.append(st)
        if stage == 0:
            abs_tol = max(gmres_abs_tol, st.norm_pre_b * gmres_rel_tol)
        u = [u[a] + du[a] for a in range(dm)]
        p = p + dp
    o.set_gmres_abs_tol(gmres_abs_tol)
    return u, p, stats[0], stats[1]
