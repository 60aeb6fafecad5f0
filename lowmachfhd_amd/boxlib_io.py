"""BoxLib plotfiles and checkpoints of the momentum state, for the reference's analysis tools.

What LowMachFHD writes (host I/O around the solve; the cell fields come from the device):

* ``write_plotfile`` (src_lowMach/write_plotfile.f90:21-539): one cell-centred plotfile
  ``plt%08d`` whose variables, for one species without charges, are rho, rho1, c1, the averaged
  velocities (average_face_to_cc), optionally the shifted velocities (shift_face_to_cc), and
  pres, in that order (:144-220); and with ``plot_stag`` one plotfile per face direction,
  ``stagx%08d`` / ``stagy%08d`` / ``stagz%08d``, holding velx / vely / velz on the nodal box
  (:281-289, :420-438).
* ``checkpoint_write`` (src_lowMach/checkpoint.f90:25-235): ``chk%08d/State`` with the cell
  components rho (nspecies), rho_sum (nspecies), rhotot, pi; ``State_edge{x,y,z}`` with umac,
  umac_sum on the faces; and ``chk%08d/Header``, the namelist ``chkpoint`` (time, dt, nlevs).

The bytes follow the BoxLib plotfile layout that FBoxLib's ``fabio_ml_multifab_write_d``
produces: a ``Header`` text file ("NavierStokes-V1.1"), ``Level_0/Cell_H`` (the VisMF header:
box array, file offsets, per-component minima and maxima) and ``Level_0/Cell_D_00000`` (one FAB
per box: an ASCII descriptor line for little-endian IEEE doubles and the box, then the data,
component by component, x fastest). FBoxLib is not part of the reference tree and no stored
plotfile exists there, so this layout is written from the published format and read back by
``read_plotfile``; its byte-for-byte equality with FBoxLib's output is not pinned.
"""
from __future__ import annotations

import os
import re

import numpy as np

from .grid import Grid

# IEEE double, little-endian byte order, in BoxLib's RealDescriptor notation
REAL_DESCRIPTOR = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"


def _ituple(v):
    return "(" + ",".join(str(int(x)) for x in v) + ")"


def _box(lo, hi, typ):
    return f"({_ituple(lo)} {_ituple(hi)} {_ituple(typ)})"


def _fortran_order(a: np.ndarray, dm: int) -> np.ndarray:
    """(z, y, x) array -> flat values with x fastest (BoxLib's order); 2-D arrays carry nz = 1."""
    return np.ascontiguousarray(a, dtype="<f8").reshape(-1)


def _write_level(dirname: str, comps, lo, hi, typ, dm: int):
    """Level_0/Cell_H and Level_0/Cell_D_00000 of one box holding ``comps``."""
    lev = os.path.join(dirname, "Level_0")
    os.makedirs(lev, exist_ok=True)
    header = f"FAB {REAL_DESCRIPTOR}{_box(lo, hi, typ)} {len(comps)}\n".encode()
    with open(os.path.join(lev, "Cell_D_00000"), "wb") as f:
        f.write(header)
        for c in comps:
            f.write(_fortran_order(c, dm).tobytes())
    mins = ",".join(f"{float(np.min(c)):.17e}" for c in comps) + ","
    maxs = ",".join(f"{float(np.max(c)):.17e}" for c in comps) + ","
    with open(os.path.join(lev, "Cell_H"), "w") as f:
        f.write("1\n0\n")
        f.write(f"{len(comps)}\n0\n")
        f.write(f"(1 0\n{_box(lo, hi, typ)}\n)\n")
        f.write("1\nFabOnDisk: Cell_D_00000 0\n\n")
        f.write(f"1,{len(comps)}\n{mins}\n\n1,{len(comps)}\n{maxs}\n")


def _check(grid: Grid, arrays, shape):
    for a in arrays:
        if tuple(np.shape(a)) != tuple(shape):
            raise ValueError(f"array of shape {np.shape(a)} where {shape} is expected")


def write_multifab(dirname: str, names, arrays, grid: Grid, time: float = 0.0, step: int = 0,
                   nodal_dir: int | None = None):
    """One single-box, single-level BoxLib plotfile directory: the cell box, or the box nodal
    in ``nodal_dir`` (a face multifab, as the stag plotfiles and State_edge hold)."""
    dm = grid.dm
    n = grid.n3[:dm]
    typ = [1 if d == nodal_dir else 0 for d in range(dm)]
    lo = [0] * dm
    hi = [n[d] - 1 + typ[d] for d in range(dm)]
    shape = grid.face_shape(nodal_dir) if nodal_dir is not None else grid.cc_shape()
    _check(grid, arrays, shape)
    if len(names) != len(arrays):
        raise ValueError("one name per array")
    os.makedirs(dirname, exist_ok=True)
    prob_lo = [grid.prob_lo[d] for d in range(dm)]
    prob_hi = [grid.prob_lo[d] + n[d] * grid.dx for d in range(dm)]
    with open(os.path.join(dirname, "Header"), "w") as f:
        f.write("NavierStokes-V1.1\n")
        f.write(f"{len(names)}\n")
        for nm in names:
            f.write(f"{nm}\n")
        f.write(f"{dm}\n{time:.17e}\n0\n")
        f.write(" ".join(f"{v:.17e}" for v in prob_lo) + "\n")
        f.write(" ".join(f"{v:.17e}" for v in prob_hi) + "\n")
        f.write("\n")  # refinement ratios: none for one level
        f.write(_box([0] * dm, [n[d] - 1 for d in range(dm)], [0] * dm) + "\n")
        f.write(f"{step}\n")
        f.write(" ".join(f"{grid.dx:.17e}" for _ in range(dm)) + "\n")
        f.write("0\n0\n")  # Cartesian coordinates, no boundary width
        f.write(f"0 1 {time:.17e}\n{step}\n")
        for d in range(dm):
            f.write(f"{prob_lo[d]:.17e} {prob_hi[d]:.17e}\n")
        f.write("Level_0/Cell\n")
    _write_level(dirname, arrays, lo, hi, typ, dm)


def read_plotfile(dirname: str) -> dict:
    """Read back a single-level plotfile directory written in the layout above."""
    with open(os.path.join(dirname, "Header")) as f:
        lines = f.read().split("\n")
    if lines[0].strip() != "NavierStokes-V1.1":
        raise ValueError("not a BoxLib plotfile")
    nv = int(lines[1])
    names = [lines[2 + i].strip() for i in range(nv)]
    k = 2 + nv
    dm = int(lines[k])
    time = float(lines[k + 1])
    prob_lo = [float(v) for v in lines[k + 3].split()]
    prob_hi = [float(v) for v in lines[k + 4].split()]
    dx = [float(v) for v in lines[k + 8].split()]
    with open(os.path.join(dirname, "Level_0", "Cell_D_00000"), "rb") as f:
        head = f.readline().decode()
        m = re.match(r"FAB \(\(8, \(64 11 52 0 1 12 0 1023\)\),\(8, \(8 7 6 5 4 3 2 1\)\)\)"
                     r"\(\(([-\d,]+)\) \(([-\d,]+)\) \(([\d,]+)\)\) (\d+)", head)
        if not m:
            raise ValueError("unsupported FAB header: " + head)
        lo = [int(v) for v in m.group(1).split(",")]
        hi = [int(v) for v in m.group(2).split(",")]
        typ = [int(v) for v in m.group(3).split(",")]
        nc = int(m.group(4))
        ext = [hi[d] - lo[d] + 1 for d in range(dm)]
        npts = int(np.prod(ext))
        data = np.frombuffer(f.read(8 * npts * nc), dtype="<f8")
    shape = tuple(reversed(ext)) if dm == 3 else (1, ext[1], ext[0])
    arrays = [data[q * npts:(q + 1) * npts].reshape(shape).copy() for q in range(nc)]
    return dict(names=names, time=time, dim=dm, prob_lo=prob_lo, prob_hi=prob_hi, dx=dx, lo=lo, hi=hi,
                nodal=typ, arrays=arrays)


def write_plotfile(base: str, solver, x_u, x_p, step: int, time: float, rho: float | np.ndarray = 1.0,
                   plot_averaged_vel: bool = True, plot_shifted_vel: bool = False, plot_stag: bool = False):
    """write_plotfile (write_plotfile.f90:21-539) for one species, no charges, isothermal:
    ``<base>/plt%08d`` with rho, rho1, c1, the velocity fields and pres (computed on the device by
    lmfhd_plot_fields), and with ``plot_stag`` the face plotfiles ``stag{x,y,z}%08d``."""
    g = solver.grid
    names, fields = solver.plot_fields(x_u, x_p, averaged=plot_averaged_vel, shifted=plot_shifted_vel, pressure=True)
    rho_a = np.broadcast_to(np.asarray(rho, dtype=np.float64), g.cc_shape()).copy()
    all_names = ["rho", "rho1", "c1"] + names
    all_fields = [rho_a, rho_a, np.ones(g.cc_shape())] + fields
    out = os.path.join(base, f"plt{step:08d}")
    write_multifab(out, all_names, all_fields, g, time, step)
    written = [out]
    if plot_stag:
        for a in range(g.dm):
            d = os.path.join(base, f"stag{'xyz'[a]}{step:08d}")
            write_multifab(d, [f"vel{'xyz'[a]}"], [np.asarray(x_u[a], dtype=np.float64)], g, time, step, nodal_dir=a)
            written.append(d)
    return written


def write_checkpoint(base: str, grid: Grid, step: int, time: float, dt: float, umac, pi,
                     rho: float | np.ndarray = 1.0, rho_sum=None, umac_sum=None):
    """checkpoint_write (checkpoint.f90:25-235) for one species without charges:
    ``<base>/chk%08d`` with State = (rho, rho_sum, rhotot, pi), State_edge{x,y,z} = (umac,
    umac_sum), and Header = the namelist chkpoint (time, dt, nlevs = 1)."""
    d = os.path.join(base, f"chk{step:08d}")
    os.makedirs(d, exist_ok=True)
    rho_a = np.broadcast_to(np.asarray(rho, dtype=np.float64), grid.cc_shape()).copy()
    rs = np.zeros(grid.cc_shape()) if rho_sum is None else np.asarray(rho_sum, dtype=np.float64)
    write_multifab(os.path.join(d, "State"), ["rho1", "rho_sum1", "rhotot", "pi"],
                   [rho_a, rs, rho_a, np.asarray(pi, dtype=np.float64)], grid, time, step)
    for a in range(grid.dm):
        us = np.zeros(grid.face_shape(a)) if umac_sum is None else np.asarray(umac_sum[a], dtype=np.float64)
        write_multifab(os.path.join(d, f"State_edge{'xyz'[a]}"), ["umac", "umac_sum"],
                       [np.asarray(umac[a], dtype=np.float64), us], grid, time, step, nodal_dir=a)
    with open(os.path.join(d, "Header"), "w") as f:
        f.write(f"&CHKPOINT\n TIME={time:.17E},\n DT={dt:.17E},\n NLEVS=1,\n /\n")
    return d


def read_checkpoint(dirname: str) -> dict:
    """checkpoint_read (checkpoint.f90:237-310) of a directory written by write_checkpoint."""
    txt = open(os.path.join(dirname, "Header")).read()
    vals = dict(re.findall(r"(\w+)\s*=\s*([-+0-9.EeDd]+)", txt))
    st = read_plotfile(os.path.join(dirname, "State"))
    dm = st["dim"]
    edges = [read_plotfile(os.path.join(dirname, f"State_edge{'xyz'[a]}")) for a in range(dm)]
    return dict(time=float(vals["TIME"].replace("D", "E")), dt=float(vals["DT"].replace("D", "E")),
                nlevs=int(vals["NLEVS"]), rho=st["arrays"][0], rho_sum=st["arrays"][1], rhotot=st["arrays"][2],
                pi=st["arrays"][3], umac=[e["arrays"][0] for e in edges], umac_sum=[e["arrays"][1] for e in edges])
