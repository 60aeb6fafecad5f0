"""The plotfile cell fields on the GPU (lmfhd_plot_fields, write_plotfile.f90:358-392):
averaged_vel = 0.5 (u(i) + u(i + e_a)) (average_face_to_cc, convert_stag.f90:76-128),
shifted_vel = u(i) (shift_face_to_cc, :914-950), pres, bit for bit against the same closed
forms in numpy; and a plotfile / checkpoint written from a GPU solve reads back to the solver's
arrays."""
import numpy as np
import pytest

from lowmachfhd_amd import boxlib_io as B
from lowmachfhd_amd.grid import NO_SLIP_WALL, Grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(16, 12, 10), (24, 20)])
def test_plot_fields_closed_forms(dims):
    import torch

    from lowmachfhd_amd import solver as S

    g = Grid(dims, 1 / dims[0])
    s = S.StokesSolver(g, visc_type=1)
    rng = np.random.default_rng(3)
    u = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    p = rng.standard_normal(g.cc_shape())
    names, f = s.plot_fields(u, p, averaged=True, shifted=True, pressure=True)
    assert names[: g.dm] == [f"averaged_vel{c}" for c in "xyz"[: g.dm]] and names[-1] == "pres"
    for a in range(g.dm):
        v = np.moveaxis(u[a], 2 - a, 0)
        avg = np.moveaxis(0.5 * (v[:-1] + v[1:]), 0, 2 - a)
        assert np.array_equal(f[a], avg)
        assert np.array_equal(f[g.dm + a], np.moveaxis(v[:-1], 0, 2 - a))
    assert np.array_equal(f[-1], p)
    # device form
    x = torch.from_numpy(s.pack(u, p)).to("cuda:0")
    out = torch.zeros(g.dm * g.ncells, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    s.plot_fields_dev(x.data_ptr(), 1, out.data_ptr())
    s.sync()
    o = out.cpu().numpy().reshape(g.dm, *g.cc_shape())
    for a in range(g.dm):
        assert np.array_equal(o[a], f[a])
    s.close()


def test_plotfile_and_checkpoint_from_a_solve(tmp_path):
    from lowmachfhd_amd import solver as S

    g = Grid((16, 16, 16), 1 / 16, bc_lo=(-1, -1, NO_SLIP_WALL), bc_hi=(-1, -1, NO_SLIP_WALL))
    s = S.StokesSolver(g, visc_type=1, gmres_rel_tol=1e-10)
    one = [np.ones(g.face_shape(a)) for a in range(3)]
    s.set_coeffs(one, np.ones(g.cc_shape()), [np.ones(g.edge_shape(e)) for e in range(3)], np.zeros(g.cc_shape()), 2.0)
    rng = np.random.default_rng(4)
    b = [rng.standard_normal(g.face_shape(a)) for a in range(3)]
    for a in range(2):  # periodic duplicates equal, wall faces zero
        v = np.moveaxis(b[a], 2 - a, 0)
        v[-1] = v[0]
    np.moveaxis(b[2], 0, 0)[0] = 0
    np.moveaxis(b[2], 0, 0)[-1] = 0
    xu, xp, st = s.gmres(b, np.zeros(g.cc_shape()))
    assert st.converged == 1
    out = B.write_plotfile(str(tmp_path), s, xu, xp, step=3, time=0.5, plot_shifted_vel=True, plot_stag=True)
    r = B.read_plotfile(out[0])
    assert r["names"] == ["rho", "rho1", "c1", "averaged_velx", "averaged_vely", "averaged_velz",
                          "shifted_velx", "shifted_vely", "shifted_velz", "pres"]
    assert np.array_equal(r["arrays"][-1], xp)
    assert np.array_equal(r["arrays"][3], 0.5 * (xu[0][:, :, :-1] + xu[0][:, :, 1:]))
    for a in range(3):
        st_ = B.read_plotfile(out[1 + a])
        assert st_["names"] == [f"vel{'xyz'[a]}"] and np.array_equal(st_["arrays"][0], xu[a])
    c = B.read_checkpoint(B.write_checkpoint(str(tmp_path), g, 3, 0.5, 0.1, xu, xp))
    assert all(np.array_equal(c["umac"][a], xu[a]) for a in range(3)) and np.array_equal(c["pi"], xp)
    s.close()
