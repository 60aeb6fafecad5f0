"""The one-workgroup coarse V-cycles staged in LDS (the default) give the same bits as the same
kernels on device memory (LMFHD_STAGE=0), for GMRES solves and for the staggered multigrid
alone: periodic and walled, 3-D and 2-D, FP64 and mixed, a grid whose whole hierarchy runs in
the kernel (8^3), and a 64^3 grid with walls.  Contexts read LMFHD_STAGE when they are built."""
import numpy as np
import pytest

from lowmachfhd_amd import problems
from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, SLIP_WALL, Grid

pytestmark = pytest.mark.gpu

WALLZ = dict(bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
WALLY = dict(bc_lo=(PERIODIC, SLIP_WALL, PERIODIC), bc_hi=(PERIODIC, NO_SLIP_WALL, PERIODIC))
CASES = [
    ("per32", lambda: Grid((32, 32, 32), 1 / 32), problems.config1_like),
    ("zwall24", lambda: Grid((24, 24, 24), 1 / 24, **WALLZ), problems.config2_like),
    ("per8", lambda: Grid((8, 8, 8), 1 / 8), problems.config1_like),
    ("per2d_64", lambda: Grid((64, 64), 1 / 64), problems.config1_like),
    ("ywall64", lambda: Grid((64, 64, 64), 1 / 64, **WALLY), problems.config2_like),
]
FORMS = [("LMFHD_STAGE", "0")]


def _solve(g, recipe, mixed, env, monkeypatch):
    from lowmachfhd_amd import solver as S

    if env:
        monkeypatch.setenv(*env)
    s = S.StokesSolver(g, visc_type=-1, precision=S.MIXED if mixed else S.FP64,
                       krylov=S.FGMRES if mixed else S.GMRES_LEFT, gmres_max_inner=30)
    if env:
        monkeypatch.delenv(env[0], raising=False)
    prob = recipe(g, seed=5, averager=s)
    s.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
    xu, xp, st = s.gmres(prob["b_u"], prob["b_p"])
    rng = np.random.default_rng(3)
    r = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
    phi = s.stag_mg_solver(r)
    s.close()
    return xu, xp, st, phi


@pytest.mark.parametrize("form", range(len(FORMS)))
@pytest.mark.parametrize("mixed", [0, 1])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_staged_coarse_matches_device_memory(case, mixed, form, monkeypatch):
    name, mk, recipe = CASES[case]
    g = mk()
    xu1, xp1, st1, phi1 = _solve(g, recipe, mixed, FORMS[form], monkeypatch)
    xu2, xp2, st2, phi2 = _solve(g, recipe, mixed, None, monkeypatch)
    print(f"{name} mixed={mixed} {FORMS[form][0]}: iterations {st1.iters} / {st2.iters}", flush=True)
    assert st1.converged == 1 and st2.iters == st1.iters
    for a, b in zip(xu1 + [xp1] + phi1, xu2 + [xp2] + phi2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mixed", [0, 1])
@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_coarse_thread_counts_same_bits(case, mixed):
    """The staggered multigrid's one-workgroup V-cycle with 256, 512 and 1024 threads: the items
    of a phase are independent (each component's loop starts on other waves, depending on the
    thread count), so the results are the same bits."""
    from lowmachfhd_amd import solver as S

    name, mk, recipe = CASES[case]
    g = mk()
    kn = dict(visc_type=-1, precision=S.MIXED if mixed else S.FP64,
              krylov=S.FGMRES if mixed else S.GMRES_LEFT, gmres_max_inner=30)
    default = S.get_tune(g, **kn)
    outs = []
    try:
        for nt in (256, 512, 1024):
            S.set_tune(g, {"coarse_threads": nt}, **kn)
            s = S.StokesSolver(g, **kn)
            assert s.tune["coarse_threads"] == nt
            prob = recipe(g, seed=5, averager=s)
            s.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
            rng = np.random.default_rng(3)
            r = [rng.standard_normal(g.face_shape(a)) for a in range(g.dm)]
            outs.append(s.stag_mg_solver(r))
            s.close()
    finally:
        S.set_tune(g, {k: default[k] for k in default if k != "ms_per_solve"}, **kn)
    print(f"{name} mixed={mixed}: threads 256 / 512 / 1024 compared", flush=True)
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
