"""Grid description and array layout shared by the HIP library and its host mirror.

Every field is held on its "valid region" for one box covering the whole domain, exactly the
index range FBoxLib calls valid for that centring (``staggered_grid/src_gmres/stag_mg_solver.f90``
builds faces with ``multifab_build_edge`` and edges with nodal flags):

* cell-centred: ``n_x * n_y * n_z`` values;
* face-centred component ``d``: ``n_d + 1`` points in direction ``d`` (both the lo face and
  the duplicated periodic / wall face at ``n_d`` are stored), ``n`` in the others;
* 3-D edges ``xy, xz, yz`` (nodal in two directions); 2-D nodes (nodal in x and y).

Arrays are x-fastest.  In numpy (C order) a field is ``(n_z, n_y, n_x)`` with the nodal
extensions, and a 2-D grid uses ``n_z = 1``.  Ghost cells are never stored: kernels map an
out-of-range index to its periodic image or wall reflection.

Boundary codes are the reference's physical codes (``src_common/bc.f90:23-50``): ``-1``
periodic, ``100..103`` no-slip walls, ``200..203`` slip walls.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

PERIODIC = -1
NO_SLIP_WALL = 100
SLIP_WALL = 200


def is_wall(code: int) -> bool:
    return 100 <= code <= 103 or 200 <= code <= 203


def is_noslip(code: int) -> bool:
    return 100 <= code <= 103


@dataclass
class Grid:
    """One box covering the domain [prob_lo, prob_hi], uniform spacing ``dx``."""

    n: tuple
    dx: float
    bc_lo: tuple = (PERIODIC, PERIODIC, PERIODIC)
    bc_hi: tuple = (PERIODIC, PERIODIC, PERIODIC)
    prob_lo: tuple = (0.0, 0.0, 0.0)
    dz2d: float = 1.0  # depth of a 2-D cell in the stochastic variance (probin_common defaults)
    dm: int = field(init=False)

    def __post_init__(self):
        n = tuple(int(v) for v in self.n)
        if len(n) not in (2, 3):
            raise ValueError("grid must be 2-D or 3-D")
        self.dm = len(n)
        self.n = n
        for d in range(self.dm):
            lo, hi = self.bc_lo[d], self.bc_hi[d]
            if (lo == PERIODIC) != (hi == PERIODIC):
                raise ValueError("periodicity must match on both sides of a direction")
            for c in (lo, hi):
                if c != PERIODIC and not is_wall(c):
                    raise ValueError(f"unsupported boundary code {c}")
        self.bc_lo = tuple(self.bc_lo) + (PERIODIC,) * (3 - len(self.bc_lo))
        self.bc_hi = tuple(self.bc_hi) + (PERIODIC,) * (3 - len(self.bc_hi))

    # ------------------------------------------------------------------ shapes (numpy C order)
    @property
    def n3(self):
        return (self.n[0], self.n[1], self.n[2] if self.dm == 3 else 1)

    @property
    def ncells(self) -> int:
        nx, ny, nz = self.n3
        return nx * ny * nz

    def cc_shape(self):
        nx, ny, nz = self.n3
        return (nz, ny, nx)

    def face_shape(self, d: int):
        ext = [0, 0, 0]
        ext[d] = 1
        nx, ny, nz = self.n3
        return (nz + ext[2], ny + ext[1], nx + ext[0])

    def edge_nodal(self, e: int):
        """Nodal directions of edge array ``e`` (3-D: 0 xy, 1 xz, 2 yz; 2-D: the node array)."""
        if self.dm == 2:
            return (0, 1)
        return ((0, 1), (0, 2), (1, 2))[e]

    def edge_shape(self, e: int):
        ext = [0, 0, 0]
        for d in self.edge_nodal(e):
            ext[d] = 1
        nx, ny, nz = self.n3
        return (nz + ext[2], ny + ext[1], nx + ext[0])

    @property
    def nedge(self) -> int:
        return 3 if self.dm == 3 else 1

    # ------------------------------------------------------------------ coordinates
    def face_coords(self, d: int):
        """Physical coordinates (X, Y, Z) of the points of face array ``d``."""
        return self._coords(tuple(i == d for i in range(3)))

    def cc_coords(self):
        return self._coords((False, False, False))

    def edge_coords(self, e: int):
        nod = self.edge_nodal(e)
        return self._coords(tuple(i in nod for i in range(3)))

    def _coords(self, nodal):
        nx, ny, nz = self.n3
        out = []
        for d, m in enumerate((nx, ny, nz)):
            if d >= self.dm:
                out.append(np.zeros(1))
                continue
            if nodal[d]:
                out.append(self.prob_lo[d] + np.arange(m + 1) * self.dx)
            else:
                out.append(self.prob_lo[d] + (np.arange(m) + 0.5) * self.dx)
        Z, Y, X = np.meshgrid(out[2], out[1], out[0], indexing="ij")
        return X, Y, Z

    def coarsen(self, f: int) -> "Grid":
        n = tuple(v // f for v in self.n)
        return Grid(n, self.dx * f, self.bc_lo[: self.dm], self.bc_hi[: self.dm], self.prob_lo, self.dz2d)

    def has_walls(self) -> bool:
        return any(is_wall(self.bc_lo[d]) or is_wall(self.bc_hi[d]) for d in range(self.dm))


def nlevs_mg(grid: Grid, minwidth: int = 2) -> int:
    """Number of multigrid levels (``src_gmres/stag_mg_layout.f90:179-222``): halve each
    direction while it is even and not yet ``minwidth``; the minimum over directions."""
    nl = None
    for d in range(grid.dm):
        t, r = grid.n[d], 1
        while t % 2 == 0 and t // minwidth != 1:
            t //= 2
            r += 1
        nl = r if nl is None else min(nl, r)
    return nl
