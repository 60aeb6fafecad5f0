"""The reference's stochastic-stress noise stream, restated on the host.

With use_bl_rng = F (probin_common.f90:316) the reference draws every Gaussian number through
HydroGrid: a 32-bit Mersenne Twister (HydroGrid/RNGs.c:62-123; seeded by srandgen, a uniform
y * 2.3283064370807969e-10 in [0,1), exact zeros redrawn) turned into N(0,1) by Acklam's
rational approximation in the Fortran form NormalRNG_dp (HydroGrid/RNG.f90:166-215).  That
form differs from the C InvNormDist (RNGs.c:180-262) in its coefficient digits and in its
upper break point, 0.9575 instead of 0.97575: for p in (0.9575, 0.97575] it takes the tail
formula.  Both are kept as written.

fill_m_stochastic (src_lowMach/stochastic_m_fluxdiv.f90:781-859) consumes the stream per RNG
stage in this order, each multifab filled whole, ghost cells included, in Fortran order (i
fastest, then j, k, component):
  * the cell stress: dm components on the box grown by one ghost cell (init_m_stochastic,
    :883-885, max(1, filtering_width) = 1); form 1 scales it by sqrt(2);
  * 3-D: the xy, xz, yz edge stresses, 2 components each, nodal in their two directions, no
    ghost cells; form 1 fills component 1 only and copies it to component 2;
    2-D: the node stress likewise.
Only the valid regions reach the flux divergence (the ghost cells are overwritten by the
periodic fill, or unused next to walls), so the arrays returned here are the valid regions.

What this cannot reproduce: which MT seed a run of the reference uses.  SeedParallelRNG
(HydroGrid/ParallelRNG.f90:19-87) derives it from the root seed through the Fortran
intrinsic RANDOM_NUMBER, whose algorithm depends on the compiler; the seed here is the value
the reference passes to SeedRNG.
"""
from __future__ import annotations

import numpy as np

from .grid import Grid

MT_SCALE = 2.3283064370807969e-10  # RNGs.c genrand

# NormalRNG_dp coefficients as written in HydroGrid/RNG.f90
_A = (-39.6968302866538, 220.946098424521, -275.928510446969, 138.357751867269, -30.6647980661472,
      2.50662827745924)
_B = (-54.4760987982241, 161.585836858041, -155.698979859887, 66.8013118877197, -13.2806815528857)
_C = (-0.00778489400243029, -0.322396458041136, -2.40075827716184, -2.54973253934373, 4.37466414146497,
      2.93816398269878)
_D = (0.00778469570904146, 0.32246712907004, 2.445134137143, 3.75440866190742)
P_LOW = 0.02425
P_HIGH = 0.9575


def acklam_fortran(p: np.ndarray) -> np.ndarray:
    """NormalRNG_dp's map from p in (0,1) to N(0,1), evaluated in the same operation order."""
    p = np.asarray(p, dtype=np.float64)
    out = np.empty_like(p)
    lo = p < P_LOW
    hi = p > P_HIGH
    mid = ~(lo | hi)
    c1, c2, c3, c4, c5, c6 = _C
    d1, d2, d3, d4 = _D
    if lo.any():
        q = np.sqrt(-2.0 * np.log(p[lo]))
        out[lo] = (((((c1 * q + c2) * q + c3) * q + c4) * q + c5) * q + c6) / \
                  ((((d1 * q + d2) * q + d3) * q + d4) * q + 1.0)
    if hi.any():
        q = np.sqrt(-2.0 * np.log(1.0 - p[hi]))
        out[hi] = -(((((c1 * q + c2) * q + c3) * q + c4) * q + c5) * q + c6) / \
                  ((((d1 * q + d2) * q + d3) * q + d4) * q + 1.0)
    if mid.any():
        a1, a2, a3, a4, a5, a6 = _A
        b1, b2, b3, b4, b5 = _B
        q = p[mid] - 0.5
        r = q * q
        out[mid] = (((((a1 * r + a2) * r + a3) * r + a4) * r + a5) * r + a6) * q / \
                   (((((b1 * r + b2) * r + b3) * r + b4) * r + b5) * r + 1.0)
    return out


class HydroGridRNG:
    """The HydroGrid MT19937 stream (srandgen(seed), then genrand / NormalRNG_dp draws).

    numpy's MT19937 with its legacy integer seeding runs the same initialisation
    (mt[i] = 1812433253 (mt[i-1] ^ (mt[i-1] >> 30)) + i) and the same tempered 32-bit outputs.
    """

    def __init__(self, seed: int):
        self._bg = np.random.RandomState(int(seed) & 0xFFFFFFFF)._bit_generator

    def raw(self, n: int) -> np.ndarray:
        return self._bg.random_raw(int(n)).astype(np.uint64)

    def uniform(self, n: int) -> np.ndarray:
        """genrand: y * 2.3283064370807969e-10."""
        return self.raw(n).astype(np.float64) * MT_SCALE

    def normals(self, n: int) -> np.ndarray:
        """n NormalRNG_dp draws: each takes the next nonzero uniform."""
        n = int(n)
        u = self.uniform(n)
        nz = u[u != 0.0]
        while nz.size < n:
            more = self.uniform(n - nz.size)
            nz = np.concatenate([nz, more[more != 0.0]])
        return acklam_fortran(nz[:n])


def _fortran_fill(rng: HydroGridRNG, ext, ncomp):
    """A multifab of extents ext (x, y, z) and ncomp components filled in Fortran order;
    returned as numpy arrays (z, y, x) per component."""
    nx, ny, nz = ext
    v = rng.normals(nx * ny * nz * ncomp)
    arr = v.reshape(ncomp, nz, ny, nx)
    return [np.ascontiguousarray(arr[c]) for c in range(ncomp)]


def fill_m_stochastic(grid: Grid, rng: HydroGridRNG, stoch_stress_form: int = 1):
    """One RNG stage of fill_m_stochastic: (W_cc, W_ed) in the layout stochastic_m_fluxdiv of
    this package takes (W_cc: dm cell arrays; W_ed: xy c1, xy c2, xz c1, xz c2, yz c1, yz c2 in
    3-D, or node c1, c2 in 2-D)."""
    dm = grid.dm
    nx, ny, nz = grid.n3
    g3 = (nx + 2, ny + 2, nz + 2 if dm == 3 else 1)
    cc = _fortran_fill(rng, g3, dm)
    if dm == 3:
        cc = [c[1:-1, 1:-1, 1:-1] for c in cc]
    else:
        cc = [c[:, 1:-1, 1:-1] for c in cc]
    if stoch_stress_form != 0:
        cc = [np.sqrt(2.0) * c for c in cc]
    W_cc = [np.ascontiguousarray(c) for c in cc]
    W_ed = []
    for e in range(grid.nedge):
        shp = grid.edge_shape(e)  # (z, y, x)
        ext = (shp[2], shp[1], shp[0])
        if stoch_stress_form == 0:
            c1, c2 = _fortran_fill(rng, ext, 2)
        else:
            (c1,) = _fortran_fill(rng, ext, 1)
            c2 = c1.copy()
        W_ed += [c1, c2]
    return W_cc, W_ed
