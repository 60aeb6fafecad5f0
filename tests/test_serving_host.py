"""Host logic of the serving loop (admission against the HBM budget, slot reuse, eviction),
driven by a stand-in solver so that it runs without a GPU."""
import types

import numpy as np
import pytest

from lowmachfhd_amd.grid import Grid
from lowmachfhd_amd.serving import Request, StokesServer, estimate_bytes, grid_key
from lowmachfhd_amd.solver import FGMRES, MIXED


class FakeTensor:
    def __init__(self, n):
        self.n = n

    def data_ptr(self):
        return 4096


class FakeTorch:
    class cuda:
        @staticmethod
        def current_stream():
            return types.SimpleNamespace(synchronize=lambda: None)

    @staticmethod
    def zeros_like(t):
        return FakeTensor(t.n)


class FakeSolver:
    """Finishes each solve after a fixed number of polls; records what ran concurrently."""
    live = []
    log = []

    def __init__(self, grid, visc_type, precision, krylov, **kn):
        self.grid = grid
        self.device_bytes = grid.ncells * 100
        self.left = 0
        self.closed = False
        FakeSolver.live.append(self)

    def set_coeffs_dev(self, ptr, theta):
        assert not self.closed

    def gmres_dev_async(self, b, x):
        assert self.left == 0 and not self.closed
        self.left = 3
        FakeSolver.log.append(sum(1 for s in FakeSolver.live if s.left > 0))

    def poll(self):
        if self.left > 0:
            self.left -= 1
        return self.left == 0

    def wait(self):
        self.left = 0
        return types.SimpleNamespace(iters=7, converged=1, true_rel_resid=1e-10)

    def close(self):
        self.closed = True
        FakeSolver.live.remove(self)


def _reqs(sizes, dt=0.0):
    out = []
    for q, n in enumerate(sizes):
        g = Grid((n, n, n), 1.0 / n)
        out.append(Request(q, q * dt, g, -1, FakeTensor(1), 20.0, FakeTensor(1)))
    return out


def test_all_served_within_max_active():
    FakeSolver.live, FakeSolver.log = [], []
    srv = StokesServer(10 ** 12, max_active=3, solver_factory=FakeSolver, torch_module=FakeTorch)
    reqs = _reqs([8, 8, 16, 8, 16, 8, 8, 16])
    res = srv.run(reqs, poll_sleep=0.0)
    assert res["n"] == 8 and res["all_converged"]
    assert all(r.finished >= r.admitted >= r.arrival for r in reqs)
    assert max(FakeSolver.log) <= 3
    # contexts are reused per grid key: at most max_active per key
    keys = [grid_key(s.grid, -1) for s in FakeSolver.live]
    assert all(keys.count(k) <= 3 for k in set(keys))


def test_budget_evicts_idle_contexts_of_other_keys():
    FakeSolver.live, FakeSolver.log = [], []
    # room for one 16^3 context (409600 B) or several 8^3 ones (51200 B each), not both kinds
    srv = StokesServer(450_000, max_active=4, solver_factory=FakeSolver, torch_module=FakeTorch)
    srv.known_bytes = {grid_key(Grid((n, n, n), 1.0 / n), -1): n ** 3 * 100 for n in (8, 16)}
    reqs = _reqs([8, 8, 16, 8, 16])
    res = srv.run(reqs, poll_sleep=0.0)
    assert res["all_converged"]
    assert res["peak_hbm_GiB"] * 2 ** 30 <= 450_000
    assert srv.contexts_built >= 3  # 8^3 contexts were released for the 16^3 ones and rebuilt


def test_request_larger_than_budget_is_rejected():
    srv = StokesServer(1000, max_active=2, solver_factory=FakeSolver, torch_module=FakeTorch)
    with pytest.raises(ValueError):
        srv.run(_reqs([16]), poll_sleep=0.0)


def test_estimate_bounds_the_footprint_order():
    g = Grid((64, 64, 64), 1 / 64)
    e = estimate_bytes(g, MIXED, FGMRES, 30)
    vlen = 3 * 65 * 64 * 64 + 64 ** 3
    assert e > 8 * 61 * vlen  # at least the Krylov vectors
    assert e < 2 * 8 * 70 * vlen


class FailingSolver(FakeSolver):
    """Every third solve fails in wait(), as lmfhd_wait reports a failed asynchronous solve."""
    started = 0
    synced = []

    def gmres_dev_async(self, b, x):
        super().gmres_dev_async(b, x)
        FailingSolver.started += 1
        self.fail = FailingSolver.started % 3 == 0

    def wait(self):
        self.left = 0
        if self.fail:
            raise RuntimeError("lmfhd error E_NAN: non-finite residual norm")
        return types.SimpleNamespace(iters=7, converged=1, true_rel_resid=1e-10)

    def sync(self):
        FailingSolver.synced.append(self)


def test_failed_solve_fails_its_request_only():
    FakeSolver.live, FakeSolver.log = [], []
    FailingSolver.started, FailingSolver.synced = 0, []
    srv = StokesServer(10 ** 12, max_active=2, solver_factory=FailingSolver, torch_module=FakeTorch)
    reqs = _reqs([8, 8, 8, 8, 8, 8, 8])
    res = srv.run(reqs, poll_sleep=0.0)
    failed = [r for r in reqs if r.error is not None]
    assert res["n"] == 7 and res["n_failed"] == len(failed) == 2
    assert all("E_NAN" in r.error and r.converged == 0 for r in failed)
    assert all(r.converged == 1 and r.error is None for r in reqs if r not in failed)
    # the failed solves' contexts were synchronised before their buffers were dropped, and the
    # slots went back to the pool: no slot is left busy
    assert len(FailingSolver.synced) == 2
    assert not any(s.busy for s in srv.slots)
    srv.close()


def test_each_size_class_tuned_once_before_its_first_context():
    FakeSolver.live, FakeSolver.log = [], []
    calls = []

    def tuner(g, **kn):
        # tuned before any context of the class exists
        assert not any(grid_key(s.grid, -1) == grid_key(g, -1) for s in FakeSolver.live)
        calls.append((tuple(g.n3), kn["visc_type"], kn["precision"], kn["krylov"]))
        return {"graphs": 1}

    srv = StokesServer(10 ** 12, max_active=3, solver_factory=FakeSolver, torch_module=FakeTorch, autotune=tuner)
    res = srv.run(_reqs([8, 16, 8, 8, 16, 8]), poll_sleep=0.0)
    assert res["all_converged"] and res["classes_tuned"] == 2
    assert sorted(calls) == [((8, 8, 8), -1, MIXED, FGMRES), ((16, 16, 16), -1, MIXED, FGMRES)]
    srv.tune_class(Grid((8, 8, 8), 1.0 / 8), -1)  # already tuned: not again
    assert len(calls) == 2
