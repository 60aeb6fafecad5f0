"""The C ABI library loads (no device call) and exports every entry point include/lmfhd.h declares."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "lmfhd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lmfhd_\w+)\s*\(", src)))


def test_header_symbols_exported():
    from lowmachfhd_amd import solver

    lib = ctypes.CDLL(solver.LIB_PATH)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(solver.EXPORTS)


def test_desc_defaults_mirror_probin():
    """lmfhd_desc_defaults = probin_gmres.f90:122-163 / probin_common.f90 defaults (no device call)."""
    from lowmachfhd_amd import solver

    L = solver.load()
    d = solver.Desc()
    L.lmfhd_desc_defaults(ctypes.byref(d))
    assert (d.precon_type, d.gmres_max_inner, d.gmres_max_outer, d.gmres_max_iter, d.gmres_min_iter) == (1, 5, 20, 100, 1)
    assert d.gmres_rel_tol == 1e-9 and d.gmres_abs_tol == 0.0 and d.p_norm_weight == 1.0
    assert (d.stag_mg_nsmooths_down, d.stag_mg_nsmooths_up, d.stag_mg_nsmooths_bottom) == (2, 2, 8)
    assert (d.stag_mg_minwidth, d.stag_mg_max_vcycles, d.stag_mg_smoother, d.stag_mg_omega) == (2, 1, 1, 1.0)
    assert (d.mg_nsmooths_down, d.mg_nsmooths_up, d.mg_nsmooths_bottom, d.mg_minwidth, d.mg_max_vcycles) == (2, 2, 8, 2, 1)


def test_invalid_desc_rejected_without_device():
    from lowmachfhd_amd import solver

    L = solver.load()
    d = solver.Desc()
    L.lmfhd_desc_defaults(ctypes.byref(d))
    d.visc_type = 0
    ctx = ctypes.c_void_p()
    assert L.lmfhd_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -1
    d.visc_type = 1
    d.bc_lo[0] = 100  # walls must come in pairs with periodic matching
    d.bc_hi[0] = -1
    assert L.lmfhd_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -1


def test_grid_layout():
    from lowmachfhd_amd.grid import Grid

    g = Grid((4, 5, 6), 0.25)
    assert g.face_shape(0) == (6, 5, 5) and g.face_shape(2) == (7, 5, 4)
    assert g.edge_shape(0) == (6, 6, 5) and g.edge_shape(1) == (7, 5, 5) and g.edge_shape(2) == (7, 6, 4)
    g2 = Grid((4, 5), 0.25)
    assert g2.face_shape(1) == (1, 6, 4) and g2.edge_shape(0) == (1, 6, 5)
    X, Y, Z = g.face_coords(0)
    assert np.isclose(X[0, 0, 1], 0.25) and np.isclose(Y[0, 1, 0], 0.375)
