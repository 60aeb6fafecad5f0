"""BoxLib plotfile / checkpoint layout of lowmachfhd_amd.boxlib_io (host I/O; the layout is the
published BoxLib one, not pinned byte for byte against FBoxLib, which the reference tree does not
hold): round trips of cell and nodal boxes in 2-D and 3-D, the Header fields, the FAB descriptor,
the VisMF minima / maxima, and the checkpoint directories and namelist of checkpoint.f90."""
import os

import numpy as np
import pytest

from lowmachfhd_amd import boxlib_io as B
from lowmachfhd_amd.grid import NO_SLIP_WALL, Grid


@pytest.mark.parametrize("dims", [(8, 6, 4), (10, 6)])
@pytest.mark.parametrize("nodal", [None, 0, 1])
def test_plotfile_round_trip(tmp_path, dims, nodal):
    g = Grid(dims, 0.125, prob_lo=(0.2, -1.0, 0.5))
    rng = np.random.default_rng(1)
    shape = g.cc_shape() if nodal is None else g.face_shape(nodal)
    arrs = [rng.standard_normal(shape) for _ in range(3)]
    d = str(tmp_path / "plt00000007")
    B.write_multifab(d, ["a", "b", "c"], arrs, g, time=1.25, step=7, nodal_dir=nodal)
    r = B.read_plotfile(d)
    assert r["names"] == ["a", "b", "c"] and r["time"] == 1.25 and r["dim"] == g.dm
    assert np.allclose(r["prob_lo"], [0.2, -1.0, 0.5][: g.dm])
    assert np.allclose(r["prob_hi"], [0.2 + dims[0] * 0.125, -1.0 + dims[1] * 0.125, 0.5 + 4 * 0.125][: g.dm])
    assert r["dx"] == [0.125] * g.dm
    assert r["nodal"] == [1 if q == nodal else 0 for q in range(g.dm)]
    assert r["hi"] == [dims[q] - 1 + (q == nodal) for q in range(g.dm)]
    for x, y in zip(arrs, r["arrays"]):
        assert np.array_equal(x, y)
    # x fastest on disk: the first values of the data block are row x of the first component
    raw = open(os.path.join(d, "Level_0", "Cell_D_00000"), "rb").read()
    head, body = raw.split(b"\n", 1)
    assert head.startswith(b"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))")
    assert np.array_equal(np.frombuffer(body[: 8 * shape[2]], "<f8"), arrs[0].reshape(-1)[: shape[2]])
    cell_h = open(os.path.join(d, "Level_0", "Cell_H")).read().split("\n")
    assert cell_h[:4] == ["1", "0", "3", "0"] and "FabOnDisk: Cell_D_00000 0" in cell_h
    mins = [float(v) for v in cell_h[cell_h.index("1,3") + 1].rstrip(",").split(",")]
    assert np.allclose(mins, [a.min() for a in arrs], rtol=0, atol=0)
    hdr = open(os.path.join(d, "Header")).read().split("\n")
    assert hdr[0] == "NavierStokes-V1.1" and hdr[-2] == "Level_0/Cell"


def test_shape_is_checked(tmp_path):
    g = Grid((8, 8), 0.125)
    with pytest.raises(ValueError):
        B.write_multifab(str(tmp_path / "x"), ["a"], [np.zeros((1, 8, 9))], g)


@pytest.mark.parametrize("dims", [(8, 6, 4), (10, 6)])
def test_checkpoint_round_trip(tmp_path, dims):
    g = Grid(dims, 0.1, bc_lo=(NO_SLIP_WALL,) * len(dims), bc_hi=(NO_SLIP_WALL,) * len(dims))
    rng = np.random.default_rng(2)
    umac = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    pi = rng.standard_normal(g.cc_shape())
    d = B.write_checkpoint(str(tmp_path), g, 12, 0.375, 0.0125, umac, pi, rho=1.5)
    assert os.path.basename(d) == "chk00000012"
    subdirs = sorted(os.listdir(d))
    assert subdirs == sorted(["Header", "State"] + [f"State_edge{c}" for c in "xyz"[: g.dm]])
    r = B.read_checkpoint(d)
    assert r["time"] == 0.375 and r["dt"] == 0.0125 and r["nlevs"] == 1
    assert np.array_equal(r["pi"], pi) and np.all(r["rho"] == 1.5) and np.all(r["rhotot"] == 1.5)
    for a in range(g.dm):
        assert np.array_equal(r["umac"][a], umac[a]) and not r["umac_sum"][a].any()


class _StubSolver:
    """Stands in for StokesSolver.plot_fields (the device computes these fields; here the same
    closed forms in numpy) so that the writer's variable list and files run on the CPU."""

    def __init__(self, grid):
        self.grid = grid

    def plot_fields(self, x_u, x_p, averaged=True, shifted=False, pressure=True):
        dm = self.grid.dm
        names, out = [], []
        for on, tag in ((averaged, "averaged"), (shifted, "shifted")):
            if not on:
                continue
            for a in range(dm):
                v = np.moveaxis(x_u[a], 2 - a, 0)
                f = 0.5 * (v[:-1] + v[1:]) if tag == "averaged" else v[:-1]
                names.append(f"{tag}_vel{'xyz'[a]}")
                out.append(np.ascontiguousarray(np.moveaxis(f, 0, 2 - a)))
        if pressure:
            names.append("pres")
            out.append(np.asarray(x_p, dtype=np.float64))
        return names, out


@pytest.mark.parametrize("dims", [(8, 6, 4), (10, 6)])
def test_write_plotfile_variables_and_stag_files(tmp_path, dims):
    """write_plotfile.f90:144-220, 281-289: rho, rho1, c1, averaged (and shifted) velocities,
    pres in the cell plotfile; velx / vely / velz on the nodal boxes of stag{x,y,z}%08d."""
    g = Grid(dims, 0.125)
    rng = np.random.default_rng(5)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    p = rng.standard_normal(g.cc_shape())
    out = B.write_plotfile(str(tmp_path), _StubSolver(g), u, p, step=42, time=2.5, rho=1.25,
                           plot_shifted_vel=True, plot_stag=True)
    assert [os.path.basename(d) for d in out] == ["plt00000042"] + [f"stag{c}00000042" for c in "xyz"[: g.dm]]
    r = B.read_plotfile(out[0])
    xyz = "xyz"[: g.dm]
    assert r["names"] == ["rho", "rho1", "c1"] + [f"averaged_vel{c}" for c in xyz] + \
        [f"shifted_vel{c}" for c in xyz] + ["pres"]
    assert np.all(r["arrays"][0] == 1.25) and np.all(r["arrays"][2] == 1.0)
    assert np.array_equal(r["arrays"][-1], p) and r["time"] == 2.5
    for a in range(g.dm):
        st = B.read_plotfile(out[1 + a])
        assert st["nodal"] == [int(q == a) for q in range(g.dm)] and np.array_equal(st["arrays"][0], u[a])
