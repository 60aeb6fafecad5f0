"""Synthetic inputs of the benchmark configurations (seeded, host-generated).

Each builder returns the coefficients and right-hand side of one Stokes system
    (theta alpha - A0) u + grad p = f,   -div u = 0,
in the form the reference's gmres takes them (gmres.f90:32-47).  Cell fields carry one ghost
layer (periodic image, or the adjacent interior value at a wall) and are turned into face /
edge coefficients by ``average_cc_to_face`` / ``average_cc_to_edge`` of the object passed as
``averager`` (the HIP solver itself, or any object with the same two methods).  The stochastic
right-hand side is built by ``stochastic_m_fluxdiv`` of the object passed as ``rhs_builder``.

Noise is drawn from numpy's PCG64 generator seeded per problem, in a fixed order: W_cc
components 1..dm, then the edge (3-D: xy c1, xy c2, xz c1, xz c2, yz c1, yz c2) or node (2-D:
c1, c2) arrays, each N(0,1) on its full valid region.  Both the GPU and the oracle receive the
same arrays, so the stochastic forcing is identical in both.
"""
from __future__ import annotations

import numpy as np

from .grid import NO_SLIP_WALL, PERIODIC, Grid


def ghosted(grid: Grid, cc: np.ndarray) -> np.ndarray:
    """Pad a cell array with one ghost layer: periodic wrap, or edge copy at walls."""
    pads, modes = [], []
    for ax, d in enumerate((2, 1, 0)):
        if d >= grid.dm:
            pads.append((0, 0))
            modes.append(None)
        else:
            pads.append((1, 1))
            modes.append("wrap" if grid.bc_lo[d] == PERIODIC else "edge")
    out = cc
    for ax, (p, mode) in enumerate(zip(pads, modes)):
        if mode is None:
            continue
        width = [(0, 0)] * 3
        width[ax] = p
        out = np.pad(out, width, mode=mode)
    return np.ascontiguousarray(out)


def noise(grid: Grid, rng: np.random.Generator):
    W_cc = [rng.standard_normal(grid.cc_shape()) for _ in range(grid.dm)]
    W_ed = [rng.standard_normal(grid.edge_shape(q // 2)) for q in range(2 * grid.nedge)]
    return W_cc, W_ed


def fourier_field(grid: Grid, coords, rng: np.random.Generator, nmodes: int, kmax: int):
    """sum_m a_m e_m sin(2 pi k_m . x + phi_m) building blocks: returns (k, phase, dirs)."""
    k = rng.integers(1, kmax + 1, size=(nmodes, 3)) * rng.choice([-1, 1], size=(nmodes, 3))
    if grid.dm == 2:
        k[:, 2] = 0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=nmodes)
    dirs = rng.standard_normal((nmodes, 3))
    if grid.dm == 2:
        dirs[:, 2] = 0.0
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return k, phase, dirs


def _mode_sum(X, Y, Z, k, phase, w):
    out = np.zeros_like(X)
    for m in range(len(phase)):
        out += w[m] * np.sin(2 * np.pi * (k[m, 0] * X + k[m, 1] * Y + k[m, 2] * Z) + phase[m])
    return out


def body_force(grid: Grid, rng: np.random.Generator, nmodes=8, kmax=4, amp=0.1):
    """Deterministic forcing: amp * e_m sin(2 pi k_m.x + phi_m) summed over nmodes, at face centres."""
    k, phase, dirs = fourier_field(grid, None, rng, nmodes, kmax)
    f = []
    for a in range(grid.dm):
        X, Y, Z = grid.face_coords(a)
        f.append(_mode_sum(X, Y, Z, k, phase, amp * dirs[:, a]))
    return f


def _zero_wall_faces(grid: Grid, u):
    for a in range(grid.dm):
        ax = 2 - a
        for side, code in ((0, grid.bc_lo[a]), (-1, grid.bc_hi[a])):
            if code != PERIODIC:
                idx = [slice(None)] * 3
                idx[ax] = side
                u[a][tuple(idx)] = 0.0
    return u


def assemble(grid, visc_type, rho_cc, eta_cc, theta, dt, seed, averager, rhs_builder, extra_force=None, kT=1.0,
             noise_source="pcg64", stoch_stress_form=0):
    """Coefficients from cell fields and the stochastic (+ deterministic) right-hand side.
    noise_source "pcg64": numpy PCG64 seeded with `seed`, i.i.d. N(0,1) (the stress form 0 of
    the benchmark text); "hydrogrid": the reference's own stream (HydroGrid MT19937 seeded with
    `seed`, fill_m_stochastic order and stress form `stoch_stress_form`)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if noise_source == "hydrogrid":
        from .noise import HydroGridRNG, fill_m_stochastic

        W_cc, W_ed = fill_m_stochastic(grid, HydroGridRNG(seed), stoch_stress_form)
    elif noise_source == "pcg64":
        W_cc, W_ed = noise(grid, rng)
    elif noise_source == "none":  # deterministic part only (noise added later, e.g. on the device)
        W_cc = [np.zeros(grid.cc_shape()) for _ in range(grid.dm)]
        W_ed = [np.zeros(grid.edge_shape(q // 2)) for q in range(2 * grid.nedge)]
    else:
        raise ValueError(noise_source)
    alpha_fc = averager.average_cc_to_face(ghosted(grid, rho_cc))
    eta_ed = averager.average_cc_to_edge(ghosted(grid, eta_cc))
    beta_cc = np.ascontiguousarray(eta_cc)
    gamma_cc = np.zeros(grid.cc_shape())
    b_u = rhs_builder.stochastic_m_fluxdiv(W_cc, W_ed, eta_cc, eta_ed, dt, kT, 1.0)
    if extra_force is not None:
        f = extra_force(rng)
        b_u = [b + x for b, x in zip(b_u, f)]
    b_u = _zero_wall_faces(grid, b_u)
    return dict(grid=grid, visc_type=visc_type, alpha_fc=alpha_fc, beta_cc=beta_cc, beta_ed=eta_ed,
                gamma_cc=gamma_cc, theta=theta, b_u=b_u, b_p=np.zeros(grid.cc_shape()), dt=dt, seed=seed)


def config0(averager, rhs_builder, seed=10):
    """2-D 64^2 periodic, rho = eta = 1, visc_type 1, dt = 0.1, stochastic RHS."""
    g = Grid((64, 64), 1.0 / 64)
    one = np.ones(g.cc_shape())
    return assemble(g, 1, one, one, 1.0 / 0.1, 0.1, seed, averager, rhs_builder)


def config1_like(grid: Grid, seed=11, averager=None, rhs_builder=None, dt=0.05, noise_source="pcg64"):
    """Config-1 recipe on any periodic grid: eta = 1 + sin sin sin / 2, rho = 1, visc_type -1,
    stochastic RHS plus 8 random Fourier modes of amplitude 0.1."""
    rhs_builder = rhs_builder or averager
    X, Y, Z = grid.cc_coords()
    s = lambda v: np.sin(2 * np.pi * v)  # noqa: E731
    eta = 1.0 + 0.5 * s(X) * s(Y) * (s(Z) if grid.dm == 3 else 1.0)
    rho = np.ones(grid.cc_shape())
    return assemble(grid, -1, rho, eta, 1.0 / dt, dt, seed, averager, rhs_builder,
                    extra_force=lambda rng: body_force(grid, rng), noise_source=noise_source)


def config1(averager, rhs_builder, n=128, seed=11):
    return config1_like(Grid((n, n, n), 1.0 / n), seed, averager, rhs_builder)


def config2_like(grid: Grid, seed=12, averager=None, rhs_builder=None, dt=0.1, noise_source="pcg64"):
    """Config-2 recipe: c = 16 random Fourier modes clipped to [0.05, 0.95], rho = c + 2(1-c),
    eta = 1 + c, visc_type -1, stochastic RHS; walls as given by the grid."""
    rhs_builder = rhs_builder or averager
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    k, phase, _ = fourier_field(grid, None, rng, 16, 4)
    amp = rng.uniform(0.05, 0.15, size=16)
    X, Y, Z = grid.cc_coords()
    c = np.clip(0.5 + _mode_sum(X, Y, Z, k, phase, amp), 0.05, 0.95)
    rho = c + 2.0 * (1.0 - c)
    eta = 1.0 + c
    return assemble(grid, -1, rho, eta, 1.0 / dt, dt, seed, averager, rhs_builder, noise_source=noise_source)


def config2(averager, rhs_builder, n=256, seed=12):
    g = Grid((n, n, n), 1.0 / n, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    return config2_like(g, seed, averager, rhs_builder)


def config3_requests(seed=13):
    """96 independent periodic realisations: 24 each of 32^3, 48^3, 64^3, 96^3, own seeds."""
    sizes = [32] * 24 + [48] * 24 + [64] * 24 + [96] * 24
    rng = np.random.Generator(np.random.PCG64(seed))
    order = rng.permutation(len(sizes))
    return [(sizes[i], seed * 1000 + int(i)) for i in order]


def config3_problem(n, seed, averager, rhs_builder):
    g = Grid((n, n, n), 1.0 / n)
    one = np.ones(g.cc_shape())
    return assemble(g, 1, one, one, 1.0 / 0.1, 0.1, seed, averager, rhs_builder)


def config4_requests(nreq=40, rate=10.0, seed=14):
    """Serving stream: Poisson arrivals at `rate` per second; sizes uniform over
    {32, 64, 96, 128, 192}^3, boundary conditions uniform over {periodic, no-slip z walls};
    periodic requests take config-1 coefficients, walled ones config-2 coefficients.
    Returns [(arrival_s, n, walls, seed)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    t -= t[0]
    sizes = rng.choice([32, 64, 96, 128, 192], size=nreq)
    walls = rng.integers(0, 2, size=nreq)
    return [(float(t[q]), int(sizes[q]), bool(walls[q]), seed * 1000 + q) for q in range(nreq)]


def config4_grid(n, walls):
    if walls:
        return Grid((n, n, n), 1.0 / n, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    return Grid((n, n, n), 1.0 / n)


def config4_problem(n, walls, seed, averager, rhs_builder=None, noise_source="pcg64"):
    g = config4_grid(n, walls)
    if walls:
        return config2_like(g, seed, averager, rhs_builder, noise_source=noise_source)
    return config1_like(g, seed, averager, rhs_builder, noise_source=noise_source)
