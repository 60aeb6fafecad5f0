"""Benchmark: implicit staggered Stokes solves/s on one MI355X (128^3, rtol 1e-9).

Workload (one "step" = one complete Stokes solve from x = 0 on device-resident inputs):
  3-D 128^3 periodic, dx = 1/128, visc_type -1 with eta = 1 + sin sin sin / 2, rho = 1,
  theta = 1/dt = 20, right-hand side = stochastic momentum-flux divergence (form 0, dt = 0.05,
  seeded) + 8 random Fourier modes of amplitude 0.1; GMRES(30) preconditioned by the staggered
  velocity multigrid (colour GS 2/2/8) and the cell-centred MAC multigrid, rtol 1e-9.
Two forms are timed: the FP64 form (reference arithmetic, left-preconditioned GMRES of
gmres.f90) and the mixed form (FP32 multigrid inside FP64 flexible GMRES); `value` is the
mixed form.  The roofline entry is the colour Gauss-Seidel kernel of the finest level (the
kernel with the largest share of solve time; one launch smooths one parity of all three
velocity components, the level's arrays in parity-split storage, one thread per position of
the nodal box), timed with HIP events on the solver's stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def sig(v, digits=5):
    """Round floats to significant digits (keeps small residuals visible)."""
    if isinstance(v, float):
        return float(f"{v:.{digits}g}")
    return v


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=128)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "gs_color_traffic.json"))
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run device copies of 256 MiB (reference kernel for HBM byte counters)")
    p.add_argument("--batch", type=int, default=4, help="concurrent 128^3 mixed solves for the batched figure")
    p.add_argument("--no-ensemble", action="store_true")
    p.add_argument("--no-serving", action="store_true")
    p.add_argument("--no-timestep", action="store_true")
    p.add_argument("--max-active", type=int, default=8)
    p.add_argument("--serve-requests", type=int, default=40)
    p.add_argument("--serve-rate", type=float, default=10.0)
    p.add_argument("--serve-rates", default="20,40,80,160",
                   help="further Poisson rates of the serving sweep (solves/s), to find where p99 breaks")
    p.add_argument("--cpu-solves", type=int, default=5, help="timed CPU baseline solves after one warm-up")
    p.add_argument("--hbm-budget-gib", type=float, default=96.0)
    p.add_argument("--no-large", action="store_true",
                   help="skip the kernel timings at 256^3 (working set outside the Infinity Cache)")
    p.add_argument("--no-probe", action="store_true", help="skip the streaming-bandwidth probe")
    return p.parse_args()


def kernel_ms(torch, stream, fn, reps):
    """Average duration (ms) of fn() over reps calls, HIP events on the solver's stream."""
    ext = torch.cuda.ExternalStream(stream)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(ext)
    for _ in range(reps):
        fn()
    e1.record(ext)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def time_solves(torch, s, b, x, steps, warmup):
    for _ in range(warmup):
        x.zero_()
        torch.cuda.synchronize()
        s.gmres_dev(b.data_ptr(), x.data_ptr())
    torch.cuda.synchronize()
    lat = []
    stats = None
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        x.zero_()
        torch.cuda.synchronize()
        stats = s.gmres_dev(b.data_ptr(), x.data_ptr())
        s.sync()
        lat.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    return total, lat, stats


def cpu_baseline(prob, grid, threads, nsolves):
    """The FP64 oracle (C++ restatement of the reference, OpenMP) on the same problem on the host
    cores: one warm-up solve, then the median of nsolves timed solves.  Returns (median seconds,
    statistics, the solution packed like a device Stokes vector)."""
    import oracle

    o = oracle.Oracle(grid, visc_type=-1, gmres_max_inner=30, omp_threads=threads)
    o.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
    times = []
    for q in range(nsolves + 1):
        t0 = time.perf_counter()
        xu, xp, st, rc = o.gmres(prob["b_u"], prob["b_p"])
        if q > 0:
            times.append(time.perf_counter() - t0)
    x = np.concatenate([np.ravel(a) for a in xu] + [np.ravel(xp)])
    return float(np.median(times)), st, x


def _device_requests(torch, specs, make_problem, visc_type):
    """Generate problems on the host (one FP64 helper context per grid for the averaging and
    the stochastic RHS), pack them and upload them: requests are device-resident."""
    from lowmachfhd_amd.serving import Request, grid_key, pack_problem
    from lowmachfhd_amd.solver import StokesSolver

    helpers, reqs = {}, []
    for rid, (arrival, grid, seed, extra) in enumerate(specs):
        key = grid_key(grid, visc_type)
        if key not in helpers:
            helpers[key] = StokesSolver(grid, visc_type=visc_type, gmres_max_inner=1)
        prob = make_problem(grid, seed, helpers[key], extra)
        coef, b = pack_problem(prob)
        reqs.append(Request(rid, arrival, grid, prob["visc_type"], torch.from_numpy(coef).to("cuda:0"),
                            float(prob["theta"]), torch.from_numpy(b).to("cuda:0")))
    for h in helpers.values():
        h.close()
    torch.cuda.synchronize()
    return reqs


def bench_timestep(torch, g, prob, x0, steps, warmup):
    """Momentum time steps (lmfhd_advance_inertial_dev) on the 128^3 problem: rho = 1, the
    problem's eta halved for the Crank-Nicolson solves, the state starting from a solve's
    solution, dt at advective CFL 0.5 of that state, a fresh on-device stochastic forcing each
    step with k_B T set so that the thermal velocity sqrt(k_B T / (rho h^3)) is 1% of h / dt
    (explicit advection is stable only below CFL ~1, and k_B T = 1 on h^3 cells would drive
    velocities of order 10^4)."""
    from lowmachfhd_amd.solver import FGMRES, FP64, GMRES_LEFT, MIXED, StokesSolver

    h = g.dx
    nf = sum(int(np.prod(g.face_shape(a))) for a in range(g.dm))
    umax = float(np.abs(x0[:nf]).max())
    dt = min(1.0 / prob["theta"], 0.5 * h / max(umax, 1e-12))
    kT = (0.01 * h / dt) ** 2 * h ** 3
    out = {"dt": dt, "cfl": umax * dt / h, "k_B_T": kT}
    for name, prec, kry in (("fp64", FP64, GMRES_LEFT), ("mixed", MIXED, FGMRES)):
        s = StokesSolver(g, visc_type=-1, precision=prec, krylov=kry, gmres_max_inner=30)
        s.set_coeffs(prob["alpha_fc"], 0.5 * prob["beta_cc"], [0.5 * e for e in prob["beta_ed"]],
                     0.5 * prob["gamma_cc"], 1.0 / dt)
        x = torch.from_numpy(x0).to("cuda:0")
        f = torch.zeros_like(x)
        # the noise amplitude takes the full eta (the context holds eta / 2)
        eta_cc = torch.from_numpy(np.ascontiguousarray(prob["beta_cc"]).ravel()).to("cuda:0")
        eta_ed = torch.from_numpy(np.concatenate([np.ravel(e) for e in prob["beta_ed"]])).to("cuda:0")
        its = []

        def step(q):
            s.stochastic_m_fluxdiv_dev(1234, q, f.data_ptr(), dt, k_B=kT, eta_cc_ptr=eta_cc.data_ptr(),
                                       eta_ed_ptr=eta_ed.data_ptr())
            a, b = s.advance_inertial_dev(dt, x.data_ptr(), f.data_ptr())
            its.append((int(a.iters), int(b.iters)))

        for q in range(warmup):
            step(q)
        torch.cuda.synchronize()
        its.clear()
        t0 = time.perf_counter()
        for q in range(steps):
            step(warmup + q)
        torch.cuda.synchronize()
        dt_wall = time.perf_counter() - t0
        out[name] = dict(ms_per_step=1e3 * dt_wall / steps, steps_per_s=steps / dt_wall,
                         iters_pred=float(np.mean([a for a, _ in its])), iters_corr=float(np.mean([b for _, b in its])),
                         finite=bool(torch.isfinite(x).all()))
        s.close()
    return out


def bench_ensemble(torch, max_active, budget_gib, batch=0, tune=False):
    """96 independent periodic realisations, 24 each of 32^3/48^3/64^3/96^3, all submitted at
    once; FP32 multigrid inside FP64 FGMRES, rtol 1e-9.  batch = 1: the preconditioners of the
    solves of one size class in flight run in lock step as batched launches.  tune: each size
    class autotuned once, in the untimed warm-up pass."""
    from lowmachfhd_amd import problems
    from lowmachfhd_amd.grid import Grid
    from lowmachfhd_amd.serving import StokesServer

    specs = [(0.0, Grid((n, n, n), 1.0 / n), seed, None) for n, seed in problems.config3_requests()]
    reqs = _device_requests(torch, specs, lambda g, seed, h, _: problems.config3_problem(g.n3[0], seed, h, h), 1)
    srv = StokesServer(int(budget_gib * 2 ** 30), max_active=max_active, batch=batch, autotune=tune)
    srv.run(reqs)  # warm-up pass: builds the contexts (and tunes the size classes)
    res = srv.run(reqs)
    cells = sum(r.grid.ncells for r in reqs)
    res["cells_per_s"] = cells / res["wall_s"]
    res["iters"] = [int(min(r.iters for r in reqs)), int(max(r.iters for r in reqs))]
    res["max_active"] = max_active
    res["batch"] = batch
    srv.close()
    return res


def bench_serving(torch, nreq, rate, max_active, budget_gib, batch=0, tune=False):
    """Poisson arrivals of solves with sizes in {32,64,96,128,192}^3, periodic or z-walled,
    config-1/config-2 coefficients, each with its own on-device stochastic forcing; FP32
    multigrid inside FP64 FGMRES, rtol 1e-9."""
    from lowmachfhd_amd import problems
    from lowmachfhd_amd.serving import StokesServer

    stream = problems.config4_requests(nreq, rate)
    specs = [(t, problems.config4_grid(n, w), seed, w) for t, n, w, seed in stream]
    cache = {}

    def make(g, seed, h, walls):
        # coefficients and the deterministic forcing are generated once per (size, walls)
        # class; each request's stochastic forcing is generated on the device from its seed
        key = (g.n3[0], walls)
        if key not in cache:
            cache[key] = problems.config4_problem(g.n3[0], walls, seed, h, h, noise_source="none")
        return cache[key]

    reqs = _device_requests(torch, specs, make, -1)
    for r, (t, n, w, seed) in zip(reqs, stream):
        r.noise_seed = seed
        r.dt = 0.1 if w else 0.05
    srv = StokesServer(int(budget_gib * 2 ** 30), max_active=max_active, batch=batch, autotune=tune)
    srv.run(reqs)  # warm-up pass: builds the contexts of every size class (and tunes them)
    res = srv.run(reqs)
    res.update(rate_per_s=rate, max_active=max_active, classes=len(cache), batch=batch)
    srv.close()
    return res


def kernels_large(torch, n):
    """The finest-level Gauss-Seidel launch (FP32, one parity of all components) and stag_applyop
    (FP64) at n^3, where their working sets (0.9 and 1.7 GB) cannot stay in the 256 MB Infinity
    Cache.  Random positive coefficients; HIP-event times on the solver's stream."""
    from lowmachfhd_amd.grid import Grid
    from lowmachfhd_amd.solver import FGMRES, FP64, GMRES_LEFT, MIXED, StokesSolver

    g = Grid((n, n, n), 1.0 / n)
    rng = np.random.default_rng(3)
    coef = dict(alpha_fc=[1.0 + rng.random(g.face_shape(a)) for a in range(3)],
                beta_cc=1.0 + rng.random(g.cc_shape()),
                beta_ed=[1.0 + rng.random(g.edge_shape(e)) for e in range(3)],
                gamma_cc=np.zeros(g.cc_shape()))
    out = {"n": n}
    for name, prec, kry in (("fp64", FP64, GMRES_LEFT), ("mixed", MIXED, FGMRES)):
        s = StokesSolver(g, visc_type=-1, precision=prec, krylov=kry, gmres_max_inner=30)
        s.set_coeffs(coef["alpha_fc"], coef["beta_cc"], coef["beta_ed"], coef["gamma_cc"], 20.0)
        gs_ms = kernel_ms(torch, s.stream, lambda: s.stag_smooth_dev(1), 10) / 2.0
        esz = 4 if prec == MIXED else 8
        gs_bytes = 11.5 * esz * n ** 3
        out[f"gs_{name}_ms_per_launch"] = sig(gs_ms)
        out[f"gs_{name}_GBps"] = round(gs_bytes / (gs_ms * 1e-3) / 1e9, 1)
        if prec == FP64:
            phi = torch.randn(int(s.vec_len), device="cuda:0", dtype=torch.float64)
            o = torch.empty_like(phi)
            ms = kernel_ms(torch, s.stream, lambda: s.stag_applyop_dev(phi.data_ptr(), o.data_ptr()), 10)
            out["applyop_fp64_ms_per_launch"] = sig(ms)
            out["applyop_fp64_GBps"] = round(104.0 * n ** 3 / (ms * 1e-3) / 1e9, 1)
            del phi, o
        s.close()
    return out


def stream_probe():
    """Streaming rates with 3 and 13 read streams (lowmachfhd_amd/stream_probe, a child process)."""
    import subprocess

    exe = os.path.join(ROOT, "lowmachfhd_amd", "stream_probe")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        return json.loads(r.stdout) if r.returncode == 0 else {"error": r.stderr.strip()[-200:]}
    except Exception as e:
        return {"error": f"{type(e).__name__}: {e}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from lowmachfhd_amd import problems
    from lowmachfhd_amd.grid import Grid
    from lowmachfhd_amd.solver import FGMRES, FP64, GMRES_LEFT, MIXED, StokesSolver

    torch.cuda.init()
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE counters scale the others.
        # 1 GiB is written between the fill of the source and each copy, so that the source
        # is not read from the 256 MB Infinity Cache
        src = torch.ones(2 ** 25, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        flush = torch.empty(2 ** 27, dtype=torch.float64, device=dev)
        for _ in range(3):
            flush.fill_(1.0)
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst, flush
    n = args.n
    g = Grid((n, n, n), 1.0 / n)

    results = {}
    prob = None
    gs = {}
    for name, prec, kry in (("fp64", FP64, GMRES_LEFT), ("mixed", MIXED, FGMRES)):
        s = StokesSolver(g, visc_type=-1, precision=prec, krylov=kry, gmres_max_inner=30)
        if prob is None:
            prob = problems.config1_like(g, seed=11, averager=s)
        s.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
        b_host = s.pack(prob["b_u"], prob["b_p"])
        b = torch.from_numpy(b_host).to(dev)
        x = torch.zeros_like(b)
        total, lat, st = time_solves(torch, s, b, x, args.steps, args.warmup)
        results[name] = dict(solves_per_s=args.steps / total, ms_per_solve=1e3 * total / args.steps,
                             p50_ms=1e3 * float(np.percentile(lat, 50)), p99_ms=1e3 * float(np.percentile(lat, 99)),
                             iters=st.iters, n_precon=st.n_precon, true_rel_resid=st.true_rel_resid,
                             converged=st.converged, device_GiB=s.device_bytes / 2 ** 30)
        results[name]["x"] = x.cpu().numpy()
        # kernel timings on this context's stream
        stream = s.stream
        reps = 20
        # one Gauss-Seidel sweep on the finest level: 2 launches (one per parity, all components;
        # visc_type -1 decouples them)
        ms_sweep = kernel_ms(torch, stream, lambda: s.stag_smooth_dev(1), reps)
        gs[name] = ms_sweep / 2.0
        # stag_applyop (one launch, all components) in FP64
        if prec == FP64:
            phi64 = torch.randn(int(s.vec_len), device=dev, dtype=torch.float64)
            out64 = torch.empty_like(phi64)
            ms64 = kernel_ms(torch, stream, lambda: s.stag_applyop_dev(phi64.data_ptr(), out64.data_ptr()), reps)
            results["applyop_fp64_ms_per_launch"] = ms64
            results["applyop_fp64_GBps"] = 104.0 * n ** 3 / (ms64 * 1e-3) / 1e9
        s.close()
        del b, x

    # the mixed form, batched: several 128^3 solves in flight on their own contexts/streams, so
    # that the latency-bound coarse levels of one overlap the bandwidth-bound levels of another
    batched = None
    if args.batch > 1:
        import numpy as _np

        from lowmachfhd_amd.solver import gmres_dev_batch

        solvers, bs, xs = [], [], []
        for q in range(args.batch):
            sq = StokesSolver(g, visc_type=-1, precision=MIXED, krylov=FGMRES, gmres_max_inner=30)
            sq.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
            bq = torch.from_numpy(sq.pack(prob["b_u"], prob["b_p"])).to(dev)
            solvers.append(sq)
            bs.append(bq)
            xs.append(torch.zeros_like(bq))
        for _ in range(max(1, args.warmup)):
            for x_ in xs:
                x_.zero_()
            torch.cuda.synchronize()
            gmres_dev_batch(solvers, [b_.data_ptr() for b_ in bs], [x_.data_ptr() for x_ in xs], args.batch)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for x_ in xs:
                x_.zero_()
            torch.cuda.synchronize()
            st_b = gmres_dev_batch(solvers, [b_.data_ptr() for b_ in bs], [x_.data_ptr() for x_ in xs], args.batch)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        same = all(torch.equal(xs[0], x_) for x_ in xs[1:])
        batched = dict(batch=args.batch, solves_per_s=args.batch * args.steps / tb,
                       ms_per_batch=1e3 * tb / args.steps, iters=int(st_b[0].iters), identical_results=bool(same))
        for sq in solvers:
            sq.close()
        del bs, xs

    # the two forms agree (error of the mixed form against the FP64 form)
    xm, x64 = results["mixed"].pop("x"), results["fp64"].pop("x")
    nu = sum(int(np.prod(g.face_shape(a))) for a in range(3))
    err_u = float(np.abs(xm[:nu] - x64[:nu]).max() / np.abs(x64[:nu]).max())
    pm, p64 = xm[nu:] - xm[nu:].mean(), x64[nu:] - x64[nu:].mean()
    err_p = float(np.abs(pm - p64).max() / np.abs(p64).max())

    # roofline of the Gauss-Seidel kernel (mixed form, FP32), algorithmic bytes per launch (each
    # distinct array once, n^3 values of 4 B each), one parity of all components: phi x3 + half
    # of rhs/alpha/phi-write x3 + beta_cc + 3 beta_ed = 11.5 arrays
    gs_bytes = 11.5 * 4 * n ** 3
    gs_ms = gs["mixed"]
    achieved = gs_bytes / (gs_ms * 1e-3) / 1e9
    kname = "SplitPos3Fn<float,-1,3,1> (finest level, parity-split storage, one parity of all components)"
    traffic = None
    if os.path.exists(args.traffic_json):
        try:
            tj = json.load(open(args.traffic_json))
            # only a measurement of this same kernel
            if tj.get("kernel", "").split("<")[0] == kname.split("<")[0]:
                traffic = tj.get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kname, "bytes_per_launch": gs_bytes,
                "avg_launch_ms": round(gs_ms, 5)}

    large = None if args.no_large else kernels_large(torch, 2 * n)
    probe = None if args.no_probe else stream_probe()
    if probe and "sum13_128^3_GBps" in probe:
        # the same launch against what a kernel reading 13 arrays (the stencil's count) can stream
        roofline["sum13_GBps"] = probe["sum13_128^3_GBps"]
        roofline["frac_of_sum13"] = round(achieved / probe["sum13_128^3_GBps"], 4)

    timestep = None
    if not args.no_timestep:
        try:
            timestep = bench_timestep(torch, g, prob, xm, min(args.steps, 5), 1)
        except Exception as e:  # reported in the line; the headline figures stand on their own
            timestep = {"error": f"{type(e).__name__}: {e}"}
    ensemble = ensemble_unbatched = None
    if not args.no_ensemble:
        ensemble = bench_ensemble(torch, max(args.max_active, 24), args.hbm_budget_gib, batch=1)
        ensemble_unbatched = bench_ensemble(torch, args.max_active, args.hbm_budget_gib, batch=0)
    serving = sweep = None
    if not args.no_serving:
        serving = bench_serving(torch, args.serve_requests, args.serve_rate, args.max_active, args.hbm_budget_gib)
        # rate sweep: where does p99 break (the rate the server can no longer keep up with)?
        sweep = [{k: sig(serving[k]) for k in ("rate_per_s", "solves_per_s", "p50_ms", "p99_ms")}]
        for r in [float(v) for v in args.serve_rates.split(",") if v]:
            sv = bench_serving(torch, args.serve_requests, r, args.max_active, args.hbm_budget_gib)
            sweep.append({k: sig(sv[k]) for k in ("rate_per_s", "solves_per_s", "p50_ms", "p99_ms")})

    cpu = None
    mixed_vs_oracle = fp64_vs_oracle = None
    if not args.no_cpu_baseline:
        # the box's CPU share (OMP_NUM_THREADS is set to it there), at most the cores present
        threads = args.cpu_threads or min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), os.cpu_count() or 1)
        dt_cpu, st_cpu, x_orc = cpu_baseline(prob, g, threads, args.cpu_solves)
        cpu = {"value": sig(1.0 / dt_cpu), "unit": "solves/s", "cores": threads, "kind": "port",
               "host_cpus": os.cpu_count(),
               "sample": f"the same {n}^3 FP64 solve by the C++ oracle (one process, OpenMP {threads} threads, "
                         f"no MPI): median of {args.cpu_solves} solves after one warm-up, {dt_cpu:.2f} s per solve, "
                         f"{st_cpu.total_iter} GMRES iterations"}

        def err(x):
            eu = float(np.abs(x[:nu] - x_orc[:nu]).max() / np.abs(x_orc[:nu]).max())
            px, po = x[nu:] - x[nu:].mean(), x_orc[nu:] - x_orc[nu:].mean()
            return {"u": eu, "p": float(np.abs(px - po).max() / np.abs(po).max())}

        mixed_vs_oracle, fp64_vs_oracle = err(xm), err(x64)

    mixed = results["mixed"]
    line = {
        "metric": "Stokes solves/s at 128^3 rtol 1e-9 (FP64 vs mixed); applyop HBM GB/s; p50/p99",
        "value": round(mixed["solves_per_s"], 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_solve"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 Krylov + f32 multigrid",
        "data": "synthetic (seeded stochastic stress + Fourier body force)",
        "config": {"workload": f"3-D {n}^3 periodic variable-viscosity Stokes solve (visc_type -1), GMRES(30) "
                               f"+ staggered/MAC multigrid, rtol 1e-9", "n": n, "dt": 0.05},
        "fp64_solves_per_s": round(results["fp64"]["solves_per_s"], 4),
        "mixed_over_fp64": round(mixed["solves_per_s"] / results["fp64"]["solves_per_s"], 3),
        "mixed_batched": ({k: sig(v) for k, v in batched.items()} | {
            "over_fp64": round(batched["solves_per_s"] / results["fp64"]["solves_per_s"], 3)}) if batched else None,
        "fp64": {k: sig(v) for k, v in results["fp64"].items()},
        "mixed": {k: sig(v) for k, v in mixed.items()},
        "mixed_vs_fp64_error": {"u": err_u, "p": err_p},
        "mixed_vs_oracle_error": mixed_vs_oracle,
        "fp64_vs_oracle_error": fp64_vs_oracle,
        "applyop_fp64_GBps": round(results["applyop_fp64_GBps"], 1),
        "applyop_fp64_ms_per_launch": sig(results["applyop_fp64_ms_per_launch"]),
        "applyop_fp64_frac_peak": round(results["applyop_fp64_GBps"] / HBM_PEAK_GBS, 4),
        "p50_ms": round(serving["p50_ms"] if serving else mixed["p50_ms"], 3),
        "p99_ms": round(serving["p99_ms"] if serving else mixed["p99_ms"], 3),
        "latency_source": "serving stream" if serving else "back-to-back 128^3 mixed solves",
        "serving": {k: sig(v) for k, v in serving.items()} if serving else None,
        "serving_rate_sweep": sweep,
        "ensemble": {k: sig(v) for k, v in ensemble.items()} if ensemble else None,
        "ensemble_unbatched": {k: sig(v) for k, v in ensemble_unbatched.items()} if ensemble_unbatched else None,
        "timestep": ({k: ({q: sig(w) for q, w in v.items()} if isinstance(v, dict) else
                          (sig(v) if isinstance(v, (int, float)) else v))
                      for k, v in timestep.items()} | {
            "workload": f"advance_timestep_inertial momentum step at {n}^3: predictor + corrector GMRES solves, "
                        "device noise each step"}) if timestep else None,
        "kernels_256": large,
        "stream_probe": probe,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
