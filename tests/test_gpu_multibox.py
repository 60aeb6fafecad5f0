"""The C ABI with the caller's multifabs split into boxes, as a Fortran host passes
dataptr(mf, box) for each box of a boxarray (gmres.f90:32-47 takes multifabs over
boxarray_maxsize boxes, main_driver.f90:185): every fab entry point, given one ghosted fab per
box (ng = 1, 2; ghost cells filled with garbage), returns bit for bit the single-box result,
leaves every ghost cell of its outputs untouched and writes a face that two boxes hold into both.
The single-box path with offset lo labels and ghost layers is checked the same way."""
import ctypes as C

import numpy as np
import pytest
from test_gpu_parity import consistent_u, random_coeffs

from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid

pytestmark = pytest.mark.gpu
SENT = 1.0e300  # ghost and output filler: must come back unchanged (ghosts) or overwritten (valid)

CASES = {
    # 8 boxes of 8^3 (the reference deck's max_grid_size 32 on 64^3, scaled down)
    "per3_8boxes": (lambda: Grid((16, 16, 16), 1 / 16), 8),
    # uneven boxes (6, 5, 5 cells) with no-slip z walls and a slip x wall
    "walls3_uneven": (lambda: Grid((16, 12, 16), 1 / 16, bc_lo=(SLIP_WALL, PERIODIC, NO_SLIP_WALL),
                                   bc_hi=(SLIP_WALL, PERIODIC, NO_SLIP_WALL)), 6),
    "ywall2": (lambda: Grid((32, 24), 1 / 32, bc_lo=(PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, NO_SLIP_WALL)), 10),
}


def _nod(g, kind):
    if kind == "c":
        return (0, 0, 0)
    if kind[0] == "f":
        a = int(kind[1])
        return tuple(int(d == a) for d in range(3))
    e = int(kind[1])
    nd = g.edge_nodal(e)
    return tuple(int(d in nd) for d in range(3))


class MF:
    """Fields of one argument (component-major) as per-box views."""

    def __init__(self, S, g, boxes, arrs, kinds, ng, fill=SENT):
        self.g, self.boxes, self.kinds, self.ng = g, boxes, kinds, ng
        self.shapes = [a.shape for a in arrs]
        parts, self.views = [], []
        for a, k in zip(arrs, kinds):
            f, v = S.box_views(g, np.ascontiguousarray(a, dtype=np.float64), boxes, _nod(g, k), ng, fill)
            parts.append(f)
            self.views.append(v)
        self.fabs = S.concat_fabs(*parts)

    def gather(self, S):
        return [S.gather_views(self.g, v, self.boxes, _nod(self.g, k), sh)
                for v, k, sh in zip(self.views, self.kinds, self.shapes)]

    def check_ghosts_and_copies(self, S, want):
        """ghost cells unchanged; every box's copy of every valid point equals the field"""
        for v, k, w in zip(self.views, self.kinds, want):
            nod = _nod(self.g, k)
            for b, (lo, hi) in enumerate(self.boxes):
                a = v[b].copy()
                gz = self.ng if self.g.dm == 3 else 0
                ext = [hi[d] - lo[d] + 1 + (nod[d] if d < self.g.dm else 0) for d in range(3)]
                valid = a[gz:gz + ext[2], self.ng:self.ng + ext[1], self.ng:self.ng + ext[0]]
                sl = tuple(slice(lo[d], lo[d] + ext[d]) for d in (2, 1, 0)) if self.g.dm == 3 else \
                    (slice(None), slice(lo[1], lo[1] + ext[1]), slice(lo[0], lo[0] + ext[0]))
                assert np.array_equal(valid, w[sl]), (k, b)
                valid[...] = SENT
                assert np.all(a == SENT), ("ghost written", k, b)


def _faces(g):
    return [f"f{a}" for a in range(g.dm)]


def _edges(g):
    return [f"e{e}" for e in range(g.nedge)]


@pytest.mark.parametrize("ng", [1, 2])
@pytest.mark.parametrize("case", list(CASES))
def test_multibox_equals_single_box(case, ng):
    from lowmachfhd_amd import solver as S

    mk, msz = CASES[case]
    g = mk()
    boxes = S.chop_boxes(g, msz)
    assert len(boxes) > 1
    rng = np.random.default_rng(23)
    al, be, bed, ga = random_coeffs(g, rng, True)
    one = S.StokesSolver(g, visc_type=-1, gmres_max_inner=30)
    mb = S.StokesSolver(g, visc_type=-1, gmres_max_inner=30, boxes=boxes)
    L = mb.L
    one.set_coeffs(al, be, bed, ga, 4.0)
    A, B, E, G = (MF(S, g, boxes, al, _faces(g), ng), MF(S, g, boxes, [be], ["c"], ng),
                  MF(S, g, boxes, bed, _edges(g), ng), MF(S, g, boxes, [ga], ["c"], ng))
    assert L.lmfhd_set_coeffs(mb.ctx, A.fabs, B.fabs, E.fabs, G.fabs, 4.0) == 0, L.lmfhd_last_error()
    u = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    p = rng.standard_normal(g.cc_shape())

    def out(kinds, shapes):
        return MF(S, g, boxes, [np.full(sh, SENT) for sh in shapes], kinds, ng)

    U, P = MF(S, g, boxes, u, _faces(g), ng), MF(S, g, boxes, [p], ["c"], ng)
    fshapes = [g.face_shape(a) for a in range(g.dm)]
    # stag_applyop (colour 0), apply_matrix, apply_precon, compute_div, compute_grad
    want = one.stag_applyop(u, 0)
    Lo = out(_faces(g), fshapes)
    assert L.lmfhd_stag_applyop(mb.ctx, U.fabs, Lo.fabs, 0) == 0
    Lo.check_ghosts_and_copies(S, want)
    wu, wp = one.apply_matrix(u, p)
    BU, BP = out(_faces(g), fshapes), out(["c"], [g.cc_shape()])
    assert L.lmfhd_apply_matrix(mb.ctx, U.fabs, P.fabs, BU.fabs, BP.fabs) == 0
    BU.check_ghosts_and_copies(S, wu)
    BP.check_ghosts_and_copies(S, [wp])
    wu, wp = one.apply_precon(u, p)
    XU, XP = out(_faces(g), fshapes), out(["c"], [g.cc_shape()])
    assert L.lmfhd_apply_precon(mb.ctx, U.fabs, P.fabs, XU.fabs, XP.fabs) == 0
    XU.check_ghosts_and_copies(S, wu)
    XP.check_ghosts_and_copies(S, [wp])
    D = out(["c"], [g.cc_shape()])
    assert L.lmfhd_compute_div(mb.ctx, U.fabs, D.fabs) == 0
    D.check_ghosts_and_copies(S, [one.compute_div(u)])
    GR = out(_faces(g), fshapes)
    assert L.lmfhd_compute_grad(mb.ctx, P.fabs, GR.fabs) == 0
    GR.check_ghosts_and_copies(S, one.compute_grad(p))
    # multigrid solvers (phi: initial guess in, solution out)
    want = one.stag_mg_solver(u)
    PH = MF(S, g, boxes, [np.zeros(sh) for sh in fshapes], _faces(g), ng)
    assert L.lmfhd_stag_mg_solver(mb.ctx, U.fabs, PH.fabs) == 0
    PH.check_ghosts_and_copies(S, want)
    M = out(["c"], [g.cc_shape()])
    assert L.lmfhd_mac_mg_solver(mb.ctx, P.fabs, M.fabs) == 0
    M.check_ghosts_and_copies(S, [one.mac_mg_solver(p)])
    # gmres (x = 0 on entry), with the statistics of the single-box solve
    xu, xp, st1 = one.gmres(u, np.zeros(g.cc_shape()))
    Z = MF(S, g, boxes, [np.zeros(g.cc_shape())], ["c"], ng)
    XU = MF(S, g, boxes, [np.zeros(sh) for sh in fshapes], _faces(g), ng)
    XP = MF(S, g, boxes, [np.zeros(g.cc_shape())], ["c"], ng)
    st = S.Stats()
    npr = C.c_double()
    assert L.lmfhd_gmres(mb.ctx, U.fabs, Z.fabs, XU.fabs, XP.fabs, C.byref(npr), C.byref(st)) == 0
    assert st.iters == st1.iters and st.converged == 1
    XU.check_ghosts_and_copies(S, xu)
    XP.check_ghosts_and_copies(S, [xp])
    # averaging from a ghosted cell field: the domain's ghost cells come from the boxes at the
    # boundary, the boxes' interior ghosts hold garbage and must not be read
    nx, ny, nz = g.n3
    ccg = 1 + rng.random((nz + 2 if g.dm == 3 else 1, ny + 2, nx + 2))
    cc_valid = ccg[1:-1, 1:-1, 1:-1] if g.dm == 3 else ccg[:, 1:-1, 1:-1]
    CC = MF(S, g, boxes, [cc_valid], ["c"], ng, fill=-7.0)
    for b, (lo, hi) in enumerate(boxes):  # domain ghost cells into the boxes' ghost layers
        v = CC.views[0][b]
        gz = ng if g.dm == 3 else 0
        for kk in range(v.shape[0]):
            for jj in range(v.shape[1]):
                for ii in range(v.shape[2]):
                    I = (lo[0] + ii - ng, lo[1] + jj - ng, (lo[2] + kk - gz) if g.dm == 3 else 0)
                    outside = any(I[d] < 0 or I[d] >= g.n[d] for d in range(g.dm))
                    near = all(-1 <= I[d] <= g.n[d] for d in range(g.dm))
                    if outside and near:
                        v[kk, jj, ii] = ccg[(I[2] + 1) if g.dm == 3 else 0, I[1] + 1, I[0] + 1]
    AF = out(_faces(g), fshapes)
    assert L.lmfhd_average_cc_to_face(mb.ctx, CC.fabs, AF.fabs) == 0
    AF.check_ghosts_and_copies(S, one.average_cc_to_face(ccg))
    AE = out(_edges(g), [g.edge_shape(e) for e in range(g.nedge)])
    assert L.lmfhd_average_cc_to_edge(mb.ctx, CC.fabs, AE.fabs) == 0
    AE.check_ghosts_and_copies(S, one.average_cc_to_edge(ccg))
    # stochastic_m_fluxdiv from given noise
    Wcc = [rng.standard_normal(g.cc_shape()) for _ in range(g.dm)]
    Wed = [rng.standard_normal(g.edge_shape(q // 2)) for q in range(2 * g.nedge)]
    eed = [1 + rng.random(g.edge_shape(e)) for e in range(g.nedge)]
    want = one.stochastic_m_fluxdiv(Wcc, Wed, be, eed, 0.1)
    WC = MF(S, g, boxes, Wcc, ["c"] * g.dm, ng)
    WE = MF(S, g, boxes, Wed, [f"e{q // 2}" for q in range(2 * g.nedge)], ng)
    EC, EE = MF(S, g, boxes, [be], ["c"], ng), MF(S, g, boxes, eed, _edges(g), ng)
    F = out(_faces(g), fshapes)
    assert L.lmfhd_stochastic_m_fluxdiv(mb.ctx, WC.fabs, WE.fabs, EC.fabs, EE.fabs, None, None, 0.1, 1.0, 1.0,
                                        F.fabs) == 0
    F.check_ghosts_and_copies(S, want)
    one.close()
    mb.close()


def test_single_box_offset_labels_and_ghosts():
    """nboxes = 0: one fab per array whose lo is any label and which carries ghost layers."""
    from lowmachfhd_amd import solver as S

    g = Grid((12, 8, 10), 1 / 12, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    rng = np.random.default_rng(4)
    al, be, bed, ga = random_coeffs(g, rng, True)
    one = S.StokesSolver(g, visc_type=-1, gmres_max_inner=30)
    lab = S.StokesSolver(g, visc_type=-1, gmres_max_inner=30)
    one.set_coeffs(al, be, bed, ga, 2.0)
    box = [((5, -3, 7), (5 + 11, -3 + 7, 7 + 9))]  # the labels: a box with lo (5, -3, 7)

    def views(arrs, kinds, ng=2):
        # box_views cuts the field at the box bounds: use bounds relative to the labels' origin,
        # then relabel the fabs
        mf = MF(S, g, [((0, 0, 0), (11, 7, 9))], arrs, kinds, ng)
        for f in mf.fabs:
            for d in range(3):
                f.lo[d] += box[0][0][d]
                f.hi[d] += box[0][0][d]
        return mf

    A, B, E, G = views(al, _faces(g)), views([be], ["c"]), views(bed, _edges(g)), views([ga], ["c"])
    L = lab.L
    assert L.lmfhd_set_coeffs(lab.ctx, A.fabs, B.fabs, E.fabs, G.fabs, 2.0) == 0
    u = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
    p = rng.standard_normal(g.cc_shape())
    wu, wp = one.apply_matrix(u, p)
    U, P = views(u, _faces(g)), views([p], ["c"])
    BU = views([np.full(g.face_shape(a), SENT) for a in range(3)], _faces(g))
    BP = views([np.full(g.cc_shape(), SENT)], ["c"])
    assert L.lmfhd_apply_matrix(lab.ctx, U.fabs, P.fabs, BU.fabs, BP.fabs) == 0
    # the views were cut with box bounds at 0: compare with the same
    for mf, w in ((BU, wu), (BP, [wp])):
        mf.boxes = [((0, 0, 0), (11, 7, 9))]
        mf.check_ghosts_and_copies(S, w)
    one.close()
    lab.close()
