"""ctypes wrapper of the FP64 CPU oracle (``oracle/oracle.cpp``).

TEST INFRASTRUCTURE ONLY: used by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker / CPU baseline.  The product package
(``lowmachfhd_amd``) never imports it.

Arrays are numpy float64 in the valid-region layout of ``lowmachfhd_amd.grid``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from lowmachfhd_amd.grid import Grid

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("dm", C.c_int), ("n", C.c_int * 3), ("dx", C.c_double), ("dz2d", C.c_double),
        ("bc_lo", C.c_int * 3), ("bc_hi", C.c_int * 3), ("visc_type", C.c_int),
        ("stag_nu_down", C.c_int), ("stag_nu_up", C.c_int), ("stag_nu_bottom", C.c_int),
        ("stag_minwidth", C.c_int), ("stag_max_vcycles", C.c_int), ("stag_smoother", C.c_int),
        ("stag_omega", C.c_double), ("stag_rel_tol", C.c_double),
        ("mac_nu_down", C.c_int), ("mac_nu_up", C.c_int), ("mac_nu_bottom", C.c_int),
        ("mac_minwidth", C.c_int), ("mac_max_vcycles", C.c_int),
        ("gmres_max_inner", C.c_int), ("gmres_max_outer", C.c_int), ("gmres_max_iter", C.c_int),
        ("gmres_min_iter", C.c_int), ("gmres_rel_tol", C.c_double), ("gmres_abs_tol", C.c_double),
        ("p_norm_weight", C.c_double), ("precon_type", C.c_int), ("omp_threads", C.c_int),
        ("gmres_spatial_order", C.c_int), ("visc_schur_approx", C.c_int),
        ("filtering_width", C.c_int),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("total_iter", C.c_int), ("outer_iter", C.c_int), ("converged", C.c_int),
        ("n_precon", C.c_int), ("n_matrix", C.c_int),
        ("norm_b", C.c_double), ("norm_pre_b", C.c_double), ("norm_init_resid", C.c_double),
        ("norm_resid", C.c_double), ("norm_resid_stokes", C.c_double),
        ("resid_hist", C.c_double * 256),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    src = os.path.join(_HERE, "oracle.cpp")
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-B" if force else "-s", "-C", _HERE], stdout=subprocess.DEVNULL)
    return _LIB


_lib = None
_REF_RNGS = os.path.join(_HERE, "_ref", "libhydrogrid_rngs.so")
_REF_TREE = "/root/reference"


def build_ref() -> str | None:
    """Compile the reference's HydroGrid/RNGs.c unchanged into _ref/ (only where the reference
    tree is present); returns the library path or None."""
    if os.path.exists(os.path.join(_REF_TREE, "HydroGrid", "RNGs.c")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"], stdout=subprocess.DEVNULL)
    return _REF_RNGS if os.path.exists(_REF_RNGS) else None


def ref_rngs():
    """ctypes handle on the compiled reference RNGs.c (srandgen, genrand, genrandn,
    InvNormDist), or None when it cannot be built here."""
    path = _REF_RNGS if os.path.exists(_REF_RNGS) else build_ref()
    if path is None:
        return None
    L = C.CDLL(path)
    L.srandgen.argtypes = [C.c_ulong]
    L.srandgen.restype = None
    L.genrand.argtypes = [C.POINTER(C.c_double)]
    L.genrand.restype = None
    L.genrandn.argtypes = [C.POINTER(C.c_double)]
    L.genrandn.restype = None
    L.InvNormDist.argtypes = [C.c_double]
    L.InvNormDist.restype = C.c_double
    return L


_libs: dict = {}


def lib(path: str | None = None):
    """ctypes handle on the oracle library (default: the one built here; ``path``: another build
    of the same source, e.g. a copy with a planted mistake in a test)."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    if path is not None and path in _libs:
        return _libs[path]
    if True:
        if path is None and not os.path.exists(_LIB):
            build()
        L = C.CDLL(path or _LIB)
        vp = C.c_void_p
        dpp = C.POINTER(C.POINTER(C.c_double))
        dp = C.POINTER(C.c_double)
        L.orc_default_params.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_create.restype = vp
        L.orc_destroy.argtypes = [vp]
        L.orc_set_coeffs.argtypes = [vp, dpp, dp, dpp, dp, C.c_double]
        L.orc_stag_applyop.argtypes = [vp, dpp, dpp, C.c_int]
        L.orc_apply_matrix.argtypes = [vp, dpp, dp, dpp, dp]
        L.orc_apply_matrix_bc.argtypes = [vp, dpp, dp, dpp, dp, dpp, dpp]
        L.orc_convert_to_homogeneous.argtypes = [vp, dpp, dp, dpp, dpp]
        L.orc_apply_precon.argtypes = [vp, dpp, dp, dpp, dp]
        L.orc_stag_mg_solve.argtypes = [vp, dpp, dpp]
        L.orc_mac_mg_solve.argtypes = [vp, dp, dp]
        L.orc_gmres.argtypes = [vp, dpp, dp, dpp, dp, C.POINTER(OrcStats)]
        L.orc_compute_div.argtypes = [vp, dpp, dp]
        L.orc_compute_grad.argtypes = [vp, dp, dpp]
        L.orc_inner_prod.argtypes = [vp, dpp, dp, dpp, dp, dp, dp]
        L.orc_stochastic_m_fluxdiv.argtypes = [vp, dpp, dpp, dp, dpp, dp, dpp, C.c_double, C.c_double,
                                               C.c_double, dpp]
        L.orc_average_cc_to_face.argtypes = [vp, dp, dpp]
        L.orc_average_cc_to_edge.argtypes = [vp, dp, dpp]
        L.orc_nlevs_mg.argtypes = [vp, C.c_int]
        L.orc_advective_m_fluxdiv.argtypes = [vp, dpp, dpp, dpp, C.c_int]
        L.orc_diffusive_m_fluxdiv.argtypes = [vp, dpp, dpp, C.c_int]
        L.orc_set_gmres_abs_tol.argtypes = [vp, C.c_double]
        L.orc_stag_smooth_colour.argtypes = [vp, dpp, dpp, C.c_int]
        L.orc_stag_restriction.argtypes = [vp, dpp, dpp, C.c_int]
        L.orc_stag_prolongation.argtypes = [vp, dpp, dpp]
        if path is None:
            _lib = L
        else:
            _libs[path] = L
    return L


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _pp(arrs):
    arr = (C.POINTER(C.c_double) * len(arrs))(*[_p(a) for a in arrs])
    return C.cast(arr, C.POINTER(C.POINTER(C.c_double))), arr


class Oracle:
    """FP64 CPU restatement of the reference Stokes solve on one grid."""

    def __init__(self, grid: Grid, visc_type: int = 1, omp_threads: int = 0, lib_path: str | None = None, **knobs):
        self.grid = grid
        L = self.L = lib(lib_path)
        p = OrcParams()
        nx, ny, nz = grid.n3
        L.orc_default_params(C.byref(p), grid.dm, nx, ny, nz, grid.dx)
        p.dz2d = grid.dz2d
        for d in range(3):
            p.bc_lo[d] = grid.bc_lo[d]
            p.bc_hi[d] = grid.bc_hi[d]
        p.visc_type = visc_type
        p.omp_threads = omp_threads
        for k, v in knobs.items():
            if not hasattr(p, k):
                raise KeyError(k)
            setattr(p, k, v)
        self.params = p
        self.ctx = L.orc_create(C.byref(p))
        self._keep = []

    def __del__(self):
        try:
            if self.ctx:
                self.L.orc_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    # -------------------------------------------------------------- allocation helpers
    def zeros_u(self):
        return [np.zeros(self.grid.face_shape(d)) for d in range(self.grid.dm)]

    def zeros_p(self):
        return np.zeros(self.grid.cc_shape())

    # -------------------------------------------------------------- API
    def set_coeffs(self, alpha_fc, beta_cc, beta_ed, gamma_cc, theta_alpha):
        a, ka = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in alpha_fc])
        e, ke = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in beta_ed])
        b = np.ascontiguousarray(beta_cc, dtype=np.float64)
        g = np.ascontiguousarray(gamma_cc, dtype=np.float64)
        self.L.orc_set_coeffs(self.ctx, a, _p(b), e, _p(g), float(theta_alpha))

    def stag_applyop(self, phi, color=0, Lphi=None):
        out = Lphi if Lphi is not None else self.zeros_u()
        pp, k1 = _pp([np.ascontiguousarray(x) for x in phi])
        op, k2 = _pp(out)
        self.L.orc_stag_applyop(self.ctx, pp, op, int(color))
        return out

    def apply_matrix(self, x_u, x_p):
        bu, bp = self.zeros_u(), self.zeros_p()
        xu, k1 = _pp([np.ascontiguousarray(x) for x in x_u])
        bo, k2 = _pp(bu)
        self.L.orc_apply_matrix(self.ctx, xu, _p(np.ascontiguousarray(x_p)), bo, _p(bp))
        return bu, bp

    def apply_matrix_bc(self, x_u, x_p, vel_bc_n, vel_bc_t):
        """apply_matrix with inhomogeneous velocity boundary values (apply_matrix.f90:24-40):
        vel_bc_n dm face arrays (read on wall faces), vel_bc_t 2 node arrays (2-D) or 6 edge
        arrays (3-D) in the reference's order, each nodal in the wall direction and the
        component."""
        bu, bp = self.zeros_u(), self.zeros_p()
        xu, k1 = _pp([np.ascontiguousarray(x) for x in x_u])
        bo, k2 = _pp(bu)
        vn, k3 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in vel_bc_n])
        vt, k4 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in vel_bc_t])
        if self.L.orc_apply_matrix_bc(self.ctx, xu, _p(np.ascontiguousarray(x_p)), bo, _p(bp), vn, vt):
            raise ValueError("inhomogeneous apply_matrix needs gmres_spatial_order 2")
        return bu, bp

    def convert_to_homogeneous(self, b_u, b_p, vel_bc_n, vel_bc_t):
        """convert_to_homogeneous (convert_to_homogeneous.f90:19): returns b_u - L phi, b_p + D phi."""
        bu = [np.array(x, dtype=np.float64) for x in b_u]
        bp = np.array(b_p, dtype=np.float64)
        bo, k1 = _pp(bu)
        vn, k3 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in vel_bc_n])
        vt, k4 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in vel_bc_t])
        self.L.orc_convert_to_homogeneous(self.ctx, bo, _p(bp), vn, vt)
        return bu, bp

    def apply_precon(self, b_u, b_p):
        xu, xp = self.zeros_u(), self.zeros_p()
        bi, k1 = _pp([np.ascontiguousarray(x) for x in b_u])
        xo, k2 = _pp(xu)
        self.L.orc_apply_precon(self.ctx, bi, _p(np.ascontiguousarray(b_p)), xo, _p(xp))
        return xu, xp

    def stag_mg_solve(self, rhs, phi0=None):
        phi = [x.copy() for x in phi0] if phi0 is not None else self.zeros_u()
        ri, k1 = _pp([np.ascontiguousarray(x) for x in rhs])
        po, k2 = _pp(phi)
        self.L.orc_stag_mg_solve(self.ctx, ri, po)
        return phi

    def mac_mg_solve(self, rhs):
        phi = self.zeros_p()
        self.L.orc_mac_mg_solve(self.ctx, _p(np.ascontiguousarray(rhs)), _p(phi))
        return phi

    def gmres(self, b_u, b_p, x_u=None, x_p=None):
        xu = [x.copy() for x in x_u] if x_u is not None else self.zeros_u()
        xp = x_p.copy() if x_p is not None else self.zeros_p()
        bi, k1 = _pp([np.ascontiguousarray(x) for x in b_u])
        xo, k2 = _pp(xu)
        st = OrcStats()
        rc = self.L.orc_gmres(self.ctx, bi, _p(np.ascontiguousarray(b_p)), xo, _p(xp), C.byref(st))
        return xu, xp, st, rc

    def advective_m_fluxdiv(self, u, rho_fc, out=None):
        """mk_advective_m_fluxdiv: -div(rho u u) on the faces (added to `out` when given)."""
        inc = out is not None
        res = [x.copy() for x in out] if inc else self.zeros_u()
        ui, k1 = _pp([np.ascontiguousarray(x) for x in u])
        ri, k2 = _pp([np.ascontiguousarray(x) for x in rho_fc])
        ro, k3 = _pp(res)
        self.L.orc_advective_m_fluxdiv(self.ctx, ui, ri, ro, int(inc))
        return res

    def diffusive_m_fluxdiv(self, u, out=None):
        """diffusive_m_fluxdiv: A0 u with the set coefficients (added to `out` when given)."""
        inc = out is not None
        res = [x.copy() for x in out] if inc else self.zeros_u()
        ui, k1 = _pp([np.ascontiguousarray(x) for x in u])
        ro, k2 = _pp(res)
        self.L.orc_diffusive_m_fluxdiv(self.ctx, ui, ro, int(inc))
        return res

    def set_gmres_abs_tol(self, tol):
        self.L.orc_set_gmres_abs_tol(self.ctx, float(tol))

    def compute_div(self, u):
        d = self.zeros_p()
        ui, k1 = _pp([np.ascontiguousarray(x) for x in u])
        self.L.orc_compute_div(self.ctx, ui, _p(d))
        return d

    def compute_grad(self, p):
        g = self.zeros_u()
        go, k1 = _pp(g)
        self.L.orc_compute_grad(self.ctx, _p(np.ascontiguousarray(p)), go)
        return g

    def inner_prod(self, u1, p1, u2, p2):
        a, k1 = _pp([np.ascontiguousarray(x) for x in u1])
        b, k2 = _pp([np.ascontiguousarray(x) for x in u2])
        ou = np.zeros(3)
        op = np.zeros(1)
        self.L.orc_inner_prod(self.ctx, a, _p(np.ascontiguousarray(p1)), b, _p(np.ascontiguousarray(p2)),
                             _p(ou), _p(op))
        return ou[: self.grid.dm], op[0]

    def stochastic_m_fluxdiv(self, W_cc, W_ed, eta_cc, eta_ed, dt, kT=1.0, variance_coef=1.0,
                             T_cc=None, T_ed=None):
        f = self.zeros_u()
        wc, k1 = _pp([np.ascontiguousarray(x) for x in W_cc])
        we, k2 = _pp([np.ascontiguousarray(x) for x in W_ed])
        ee, k3 = _pp([np.ascontiguousarray(x) for x in eta_ed])
        fo, k4 = _pp(f)
        tc = _p(np.ascontiguousarray(T_cc)) if T_cc is not None else None
        te, k5 = _pp([np.ascontiguousarray(x) for x in T_ed]) if T_ed is not None else (None, None)
        self.L.orc_stochastic_m_fluxdiv(self.ctx, wc, we, _p(np.ascontiguousarray(eta_cc)), ee, tc, te,
                                       float(dt), float(kT), float(variance_coef), fo)
        return f

    def average_cc_to_face(self, cc_g1):
        f = self.zeros_u()
        fo, k = _pp(f)
        self.L.orc_average_cc_to_face(self.ctx, _p(np.ascontiguousarray(cc_g1)), fo)
        return f

    def average_cc_to_edge(self, cc_g1):
        e = [np.zeros(self.grid.edge_shape(q)) for q in range(self.grid.nedge)]
        eo, k = _pp(e)
        self.L.orc_average_cc_to_edge(self.ctx, _p(np.ascontiguousarray(cc_g1)), eo)
        return e

    def nlevs_mg(self, minwidth=2):
        return self.L.orc_nlevs_mg(self.ctx, minwidth)

    # -------------------------------------------------------------- multigrid pieces
    def smooth_colour(self, phi, rhs, color):
        """One colour of the staggered smoother (applyop on the colour, stag_mg_update, ghost
        fill) on this grid with the level-0 coefficients; returns the new phi."""
        out = [np.array(x, dtype=np.float64) for x in phi]
        po, k1 = _pp(out)
        ri, k2 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in rhs])
        self.L.orc_stag_smooth_colour(self.ctx, po, ri, int(color))
        return out

    def coarse_grid(self):
        g = self.grid
        return Grid(tuple(n // 2 for n in g.n), 2 * g.dx, bc_lo=g.bc_lo[:g.dm], bc_hi=g.bc_hi[:g.dm])

    def restriction(self, fine, simple=False):
        """Face restriction of the V-cycle to the grid coarsened by 2 (residual form, or the
        simple mean used for alpha)."""
        gc = self.coarse_grid()
        out = [np.zeros(gc.face_shape(d)) for d in range(gc.dm)]
        fi, k1 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in fine])
        co, k2 = _pp(out)
        self.L.orc_stag_restriction(self.ctx, fi, co, int(bool(simple)))
        return out

    def prolongation(self, coarse, fine):
        """fine + P coarse (prolongation-add of the V-cycle)."""
        out = [np.array(x, dtype=np.float64) for x in fine]
        ci, k1 = _pp([np.ascontiguousarray(x, dtype=np.float64) for x in coarse])
        fo, k2 = _pp(out)
        self.L.orc_stag_prolongation(self.ctx, ci, fo)
        return out
