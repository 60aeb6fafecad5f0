"""The batched ensemble (96 realisations of 32^3/48^3/64^3/96^3) and the Poisson serving stream at
their full sizes, through the continuous-batching server, on the GPU.

  * every solve converges to a true relative residual <= 1e-9 (FP32 multigrid inside FP64
    flexible GMRES);
  * every result is bit-for-bit the result of the same request solved alone (contexts share no
    device state and every reduction has a fixed order);
  * one request per size class of the ensemble (32^3, 48^3) is checked against the FP64 oracle
    on the same inputs: solution within 1e-7 relative (the error of the mixed form is stated).
"""
import numpy as np
import pytest

from lowmachfhd_amd import problems
from lowmachfhd_amd.grid import Grid

pytestmark = pytest.mark.gpu


def _requests(torch, specs, make_problem, visc_type, keep=None):
    """Device-resident requests; `keep` collects the host problems by request id."""
    from lowmachfhd_amd.serving import Request, grid_key, pack_problem
    from lowmachfhd_amd.solver import StokesSolver

    helpers, reqs = {}, []
    for rid, (arrival, grid, seed, extra) in enumerate(specs):
        key = grid_key(grid, visc_type)
        if key not in helpers:
            helpers[key] = StokesSolver(grid, visc_type=visc_type, gmres_max_inner=1)
        prob = make_problem(grid, seed, helpers[key], extra)
        if keep is not None:
            keep[rid] = prob
        coef, b = pack_problem(prob)
        reqs.append(Request(rid, arrival, grid, prob["visc_type"], torch.from_numpy(coef).to("cuda:0"),
                            float(prob["theta"]), torch.from_numpy(b).to("cuda:0")))
    for h in helpers.values():
        h.close()
    torch.cuda.synchronize()
    return reqs


def _alone(torch, srv, reqs):
    """Each request solved by itself (one active solve at a time)."""
    out = {}
    one = srv.max_active
    srv.max_active = 1
    for r in reqs:
        a = r.arrival
        r.arrival = 0.0
        srv.run([r])
        r.arrival = a
        out[r.rid] = (r.x.clone(), r.iters)
        r.x = None
    srv.max_active = one
    torch.cuda.synchronize()
    return out


def test_config3_ensemble_96():
    import torch

    import oracle
    from lowmachfhd_amd.serving import StokesServer

    specs = [(0.0, Grid((n, n, n), 1.0 / n), seed, None) for n, seed in problems.config3_requests()]
    assert len(specs) == 96
    host = {}
    reqs = _requests(torch, specs, lambda g, seed, h, _: problems.config3_problem(g.n3[0], seed, h, h), 1, keep=host)
    srv = StokesServer(96 << 30, max_active=8, keep_results=True)
    res = srv.run(reqs)
    mix = {r.rid: (r.x.clone(), r.iters) for r in reqs}
    print(f"config 3: {res['n']} solves in {res['wall_s']:.3f} s, iterations "
          f"{min(r.iters for r in reqs)}..{max(r.iters for r in reqs)}, max true residual "
          f"{res['max_true_rel_resid']:.2e}", flush=True)
    assert res["n_failed"] == 0 and res["all_converged"] and res["max_true_rel_resid"] <= 1e-9
    alone = _alone(torch, srv, reqs)
    for rid, (x, it) in mix.items():
        xa, ia = alone[rid]
        assert ia == it and torch.equal(x, xa), rid
    # one request per small size class against the FP64 oracle
    for n in (32, 48):
        r = next(q for q in reqs if q.grid.n3[0] == n)
        prob = host[r.rid]
        o = oracle.Oracle(r.grid, visc_type=1, gmres_max_inner=30, omp_threads=16)
        o.set_coeffs(prob["alpha_fc"], prob["beta_cc"], prob["beta_ed"], prob["gamma_cc"], prob["theta"])
        xu, xp, st, rc = o.gmres(prob["b_u"], prob["b_p"])
        assert rc == 0 and st.converged == 1
        ref = np.concatenate([np.ravel(a) for a in xu] + [np.ravel(xp)])
        nf = sum(int(np.prod(r.grid.face_shape(a))) for a in range(3))
        got = mix[r.rid][0].cpu().numpy()
        eu = np.abs(got[:nf] - ref[:nf]).max() / np.abs(ref[:nf]).max()
        gp, rp = got[nf:] - got[nf:].mean(), ref[nf:] - ref[nf:].mean()
        ep = np.abs(gp - rp).max() / np.abs(rp).max()
        print(f"config 3, {n}^3 request {r.rid}: oracle {st.total_iter} iterations, mixed {mix[r.rid][1]}; "
              f"error vs oracle u {eu:.2e} p {ep:.2e}", flush=True)
        assert eu < 1e-7 and ep < 1e-7
    srv.close()


def test_config4_serving_stream():
    import torch

    from lowmachfhd_amd.serving import StokesServer

    # 8 Poisson arrivals at 20/s: 32^3 and 64^3 periodic and walled, 128^3 periodic, 96^3 z walls
    stream = problems.config4_requests(8, 20.0, seed=17)
    assert (96, True) in [(n, w) for _, n, w, _ in stream]
    specs = [(t, problems.config4_grid(n, w), seed, w) for t, n, w, seed in stream]

    def make(g, seed, h, walls):
        return problems.config4_problem(g.n3[0], walls, seed, h, h, noise_source="none")

    reqs = _requests(torch, specs, make, -1)
    for r, (t, n, w, seed) in zip(reqs, stream):
        r.noise_seed = seed  # each request's stochastic forcing generated on the device
        r.dt = 0.1 if w else 0.05
    srv = StokesServer(64 << 30, max_active=4, keep_results=True)
    res = srv.run(reqs)
    mix = {r.rid: (r.x.clone(), r.iters) for r in reqs}
    print(f"config 4: {res['n']} requests, p50 {res['p50_ms']:.1f} ms, p99 {res['p99_ms']:.1f} ms, "
          f"max true residual {res['max_true_rel_resid']:.2e}", flush=True)
    assert res["n_failed"] == 0 and res["all_converged"] and res["max_true_rel_resid"] <= 1e-9
    alone = _alone(torch, srv, reqs)
    for rid, (x, it) in mix.items():
        xa, ia = alone[rid]
        assert ia == it and torch.equal(x, xa), rid
    srv.close()
