"""Continuous-batching Stokes solve server on one GPU.

Requests (a grid, its coefficients and right-hand side) arrive over time.  The server:
  * keeps a pool of solver contexts ("slots") keyed by grid and boundary conditions, built on
    demand and cached for reuse;
  * admits requests against a fixed HBM budget: a new context is built only when its device
    footprint fits next to the contexts that exist, else idle contexts of other keys are
    released (least recently used first) or the request waits;
  * runs every admitted solve asynchronously on its context (lmfhd_gmres_dev_async), so the
    solves of different requests overlap on the device, and polls for completions;
  * optionally tunes each size class once, before its first context is built (lmfhd_autotune:
    coarse-kernel thresholds and width, reduction width, HIP-graph replay of the Arnoldi steps
    and preconditioner applications); contexts of the class then take the tuned parameters.
Results do not depend on what else is running: contexts share no device state and every
reduction in the library has a fixed order.

The reference solves one system per gmres call (src_gmres/gmres.f90:32); this module only
schedules such calls.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from .grid import Grid
from .solver import FGMRES, MIXED, StokesSolver


def grid_key(g: Grid, visc_type: int):
    return (g.dm, tuple(g.n3), tuple(g.bc_lo), tuple(g.bc_hi), float(g.dx), visc_type)


def estimate_bytes(g: Grid, precision: int, krylov: int, m: int) -> int:
    """Upper estimate of a context's device footprint (before it is built)."""
    nf = sum(int(np.prod(g.face_shape(a))) for a in range(g.dm))
    nc = g.ncells
    vlen = nf + nc
    nvec = (m + 1) + (m if krylov == FGMRES else 0) + 5
    krylov_b = 8 * nvec * vlen
    esz = 4 if precision == MIXED else 8
    ne = sum(int(np.prod(g.edge_shape(e))) for e in range(g.nedge))
    level0 = esz * (4 * nf + 2 * nc + ne) + esz * (g.dm * nf // g.dm + 3 * nc) + 8 * (nf + 2 * nc + ne)
    mg = level0 * 8 // 7  # the coarser levels add at most 1/7
    return int(1.05 * (krylov_b + mg)) + (64 << 20)


def pack_problem(prob: dict):
    """Packed host images (coefficients, right-hand side) of a problem dict of problems.py."""
    coef = np.concatenate([np.ravel(a) for a in prob["alpha_fc"]] + [np.ravel(prob["beta_cc"])] +
                          [np.ravel(e) for e in prob["beta_ed"]] + [np.ravel(prob["gamma_cc"])])
    b = np.concatenate([np.ravel(a) for a in prob["b_u"]] + [np.ravel(prob["b_p"])])
    return coef, b


@dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    grid: Grid
    visc_type: int
    coef: object              # packed coefficients, device tensor (lmfhd_set_coeffs_dev layout)
    theta: float
    b: object                 # packed right-hand side (deterministic part), device tensor
    noise_seed: int | None = None  # when set, the stochastic stress divergence is added on the
    dt: float = 0.1                # device (lmfhd_stochastic_m_fluxdiv_dev, form 0, k_B T = 1)
    # filled in by the server
    admitted: float = -1.0
    finished: float = -1.0
    iters: int = -1
    converged: int = 0
    true_rel_resid: float = float("nan")
    x: object = None          # device solution (torch tensor) when keep_results
    error: str | None = None  # the library's message when the solve failed
    extra: dict = field(default_factory=dict)


class _Slot:
    def __init__(self, key, solver: StokesSolver, nbytes: int):
        self.key = key
        self.solver = solver
        self.nbytes = nbytes
        self.busy = False
        self.last_used = 0.0
        self.req: Request | None = None
        self.b = None
        self.x = None
        self.broken = False  # a failed solve left the context unusable: dropped from the pool


class StokesServer:
    def __init__(self, hbm_budget_bytes: int, max_active: int = 8, precision: int = MIXED, krylov: int = FGMRES,
                 keep_results: bool = False, solver_factory=None, torch_module=None, autotune=False, **knobs):
        """autotune: False, True (the library's autotuner, solver.autotune, with 2 timed solves per
        candidate) or a callable tuner(grid, **solver_knobs) run once per size class."""
        if torch_module is None:
            import torch as torch_module
        self.torch = torch_module
        self.solver_factory = solver_factory or StokesSolver
        self.budget = int(hbm_budget_bytes)
        self.max_active = max_active
        self.precision = precision
        self.krylov = krylov
        self.knobs = dict(knobs)
        self.knobs.setdefault("gmres_max_inner", 30)
        self.keep_results = keep_results
        self.slots: list[_Slot] = []
        self.known_bytes: dict = {}
        self.peak_bytes = 0
        self.contexts_built = 0
        if autotune is True:
            from .solver import autotune as _lib_autotune

            self.tuner = lambda g, **kn: _lib_autotune(g, reps=2, **kn)
        else:
            self.tuner = autotune or None
        self.tuned: dict = {}  # size-class key -> the tuner's result
        self.tune_s = 0.0      # seconds spent tuning

    def tune_class(self, g: Grid, visc_type: int):
        """Tune the size class of (g, visc_type) now, once (also usable before a run)."""
        key = grid_key(g, visc_type)
        if self.tuner is None or key in self.tuned:
            return
        t0 = time.perf_counter()
        self.tuned[key] = self.tuner(g, visc_type=visc_type, precision=self.precision, krylov=self.krylov,
                                     **self.knobs)
        self.tune_s += time.perf_counter() - t0

    # ------------------------------------------------------------------ pool
    def used_bytes(self) -> int:
        return sum(s.nbytes for s in self.slots)

    def _need(self, key, g: Grid) -> int:
        return self.known_bytes.get(key) or estimate_bytes(g, self.precision, self.krylov,
                                                           self.knobs["gmres_max_inner"])

    def _acquire(self, req: Request, now: float) -> _Slot | None:
        key = grid_key(req.grid, req.visc_type)
        for s in self.slots:
            if not s.busy and s.key == key:
                return s
        need = self._need(key, req.grid)
        if need > self.budget:
            raise ValueError(f"request {req.rid} needs {need} bytes, more than the HBM budget {self.budget}")
        # release idle contexts of other keys, least recently used first, until it fits
        idle = sorted((s for s in self.slots if not s.busy), key=lambda s: s.last_used)
        while self.used_bytes() + need > self.budget and idle:
            victim = idle.pop(0)
            victim.solver.close()
            self.slots.remove(victim)
        if self.used_bytes() + need > self.budget:
            return None
        self.tune_class(req.grid, req.visc_type)
        solver = self.solver_factory(req.grid, visc_type=req.visc_type, precision=self.precision,
                                     krylov=self.krylov, **self.knobs)
        nbytes = int(solver.device_bytes)
        self.known_bytes[key] = nbytes
        slot = _Slot(key, solver, nbytes)
        self.slots.append(slot)
        self.contexts_built += 1
        self.peak_bytes = max(self.peak_bytes, self.used_bytes())
        return slot

    def _start(self, slot: _Slot, req: Request, now: float):
        s = slot.solver
        slot.b = req.b if req.noise_seed is None else req.b.clone()
        slot.x = self.torch.zeros_like(req.b)
        self.torch.cuda.current_stream().synchronize()  # x, b are written on torch's stream only
        s.set_coeffs_dev(req.coef.data_ptr(), req.theta)
        if req.noise_seed is not None:
            s.stochastic_m_fluxdiv_dev(req.noise_seed, 0, slot.b.data_ptr(), req.dt, increment=True)
        s.gmres_dev_async(slot.b.data_ptr(), slot.x.data_ptr())
        slot.busy = True
        slot.req = req
        req.admitted = now

    def _finish(self, slot: _Slot, now: float):
        req = slot.req
        try:
            st = slot.solver.wait()
        except Exception as e:  # a failed solve fails its request only
            req.error = str(e)
            st = None
            try:
                slot.solver.sync()  # no kernel of the context may still use slot.b / slot.x
            except Exception as e2:
                req.error += f"; sync: {e2}"
                slot.broken = True
        req.finished = now
        if st is not None:
            req.iters = st.iters
            req.converged = st.converged
            req.true_rel_resid = st.true_rel_resid
        if self.keep_results and st is not None:
            req.x = slot.x
        slot.busy = False
        slot.req = None
        slot.b = slot.x = None
        slot.last_used = now
        if slot.broken:
            self.slots.remove(slot)
            try:
                slot.solver.close()
            except Exception:
                pass

    # ------------------------------------------------------------------ loop
    def run(self, requests: list[Request], poll_sleep: float = 2e-4) -> dict:
        """Serve requests arriving at req.arrival (seconds from now); returns latency stats."""
        todo = collections.deque(sorted(requests, key=lambda r: r.arrival))
        pending: list[Request] = []
        t0 = time.perf_counter()
        done = 0
        while done < len(requests):
            now = time.perf_counter() - t0
            while todo and todo[0].arrival <= now:
                pending.append(todo.popleft())
            progressed = False
            for slot in list(self.slots):
                if slot.busy and slot.solver.poll():
                    self._finish(slot, time.perf_counter() - t0)
                    done += 1
                    progressed = True
            active = sum(1 for s in self.slots if s.busy)
            for req in list(pending):
                if active >= self.max_active:
                    break
                slot = self._acquire(req, now)
                if slot is None:
                    continue
                self._start(slot, req, time.perf_counter() - t0)
                pending.remove(req)
                active += 1
                progressed = True
            if not progressed:
                if not pending and not any(s.busy for s in self.slots) and todo:
                    time.sleep(max(0.0, min(todo[0].arrival - (time.perf_counter() - t0), 0.01)))
                else:
                    time.sleep(poll_sleep)
        wall = time.perf_counter() - t0
        lat = np.array([r.finished - r.arrival for r in requests])
        return dict(n=len(requests), wall_s=wall, solves_per_s=len(requests) / wall,
                    p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                    mean_ms=1e3 * float(lat.mean()), max_ms=1e3 * float(lat.max()),
                    all_converged=bool(all(r.converged == 1 for r in requests)),
                    n_failed=sum(1 for r in requests if r.error is not None),
                    max_true_rel_resid=float(max((r.true_rel_resid for r in requests if r.error is None),
                                                 default=float("nan"))),
                    contexts_built=self.contexts_built, peak_hbm_GiB=self.peak_bytes / 2 ** 30,
                    budget_GiB=self.budget / 2 ** 30, classes_tuned=len(self.tuned), tune_s=self.tune_s)

    def close(self):
        for s in self.slots:
            if s.busy:
                try:
                    s.solver.wait()
                except Exception:
                    pass
            s.solver.close()
        self.slots = []
