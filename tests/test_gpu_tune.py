"""Launch parameters of a size class on the GPU: the row-mapped smoothing and operator launches
(rows_n) give the bits of the one-face-per-lane launches; the autotuner returns a valid record
that later contexts of the class take, and a solve with it converges in the same iterations to
the same solution (to rounding: the one-workgroup coarse kernels sum in another order)."""
import numpy as np
import pytest
from test_gpu_parity import GRIDS, consistent_u, pair

from lowmachfhd_amd.grid import Grid

pytestmark = pytest.mark.gpu


def _pair_with_rows(g, vt, rows_n, precision, krylov, rng_seed):
    from lowmachfhd_amd import solver as S

    kn = dict(visc_type=vt, precision=precision, krylov=krylov, gmres_max_inner=30)
    S.set_tune(g, {"rows_n": rows_n}, **kn)
    rng = np.random.default_rng(rng_seed)
    o, s = pair(g, vt, 5.0, rng, precision=precision, krylov=krylov, gmres_max_inner=30)
    assert s.tune["rows_n"] == rows_n
    return o, s


@pytest.mark.parametrize("gname", ["per3", "zwall3", "mixwall3", "per2", "ywall2", "per32", "mixwall32"])
@pytest.mark.parametrize("vt", [1, -1, 2])
@pytest.mark.parametrize("mixed", [False, True])
def test_rows_equal_flat_launches(gname, vt, mixed):
    from lowmachfhd_amd import solver as S

    g = GRIDS[gname]()
    prec, kry = (S.MIXED, S.FGMRES) if mixed else (S.FP64, S.GMRES_LEFT)
    out = {}
    for rows_n in (0, 4):  # 4: every level of these grids with n_x >= 4 by rows
        o, s = _pair_with_rows(g, vt, rows_n, prec, kry, 17)
        rng = np.random.default_rng(3)
        phi = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
        bu = consistent_u(g, [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)])
        bp = np.zeros(g.cc_shape())
        res = [s.stag_applyop(phi, 0)]
        res.append(s.stag_mg_solver(bu))
        xu, xp, st = s.gmres(bu, bp)
        res += [xu, [xp], [np.array([st.iters])]]
        out[rows_n] = res
        if rows_n == 4 and not mixed:  # and against the oracle
            assert max(np.abs(a - b).max() for a, b in zip(res[0], o.stag_applyop(phi, 0))) <= \
                1e-13 * max(np.abs(b).max() for b in o.stag_applyop(phi, 0))
        s.close()
    for r0, r1 in zip(out[0], out[4]):
        for a, b in zip(r0, r1):
            assert np.array_equal(a, b), gname
    S.set_tune(g, {"rows_n": 0}, visc_type=vt, precision=prec, krylov=kry, gmres_max_inner=30)


def test_autotune_record_taken_by_new_contexts():
    from lowmachfhd_amd import problems
    from lowmachfhd_amd import solver as S

    g = Grid((32, 32, 32), 1 / 32)
    kn = dict(visc_type=-1, precision=S.MIXED, krylov=S.FGMRES, gmres_max_inner=30)
    default = S.get_tune(g, **kn)
    s0 = S.StokesSolver(g, **kn)
    prob = problems.config1_like(g, seed=5, averager=s0)
    args = (prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
    s0.set_coeffs(*args)
    xu0, xp0, st0 = s0.gmres(prob["b_u"], prob["b_p"])
    s0.close()
    t = S.autotune(g, reps=2, **kn)
    print(f"autotune 32^3 mixed: default {default} -> {t}", flush=True)
    assert t["ms_per_solve"] > 0 and t["coarse_threads"] in (256, 512, 1024) and t["rows_n"] >= 0
    s1 = S.StokesSolver(g, **kn)
    got = s1.tune
    assert {k: got[k] for k in got if k != "ms_per_solve"} == {k: t[k] for k in t if k != "ms_per_solve"}
    s1.set_coeffs(*args)
    xu1, xp1, st1 = s1.gmres(prob["b_u"], prob["b_p"])
    s1.close()
    assert st1.converged == 1 and st1.iters == st0.iters
    eu = max(np.abs(a - b).max() for a, b in zip(xu1, xu0)) / max(np.abs(b).max() for b in xu0)
    assert eu < 1e-9
    S.set_tune(g, {k: default[k] for k in default if k != "ms_per_solve"}, **kn)
