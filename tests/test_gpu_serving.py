"""Serving loop on the GPU: every request's result is the same, bit for bit,
whether it is solved alone or in a mix of concurrent requests of different
sizes admitted under a memory budget that forces queueing.  Requests solved
together as one forest (batching of small requests) equal their solves
alone to within the engine's tolerance: NLL 1e-9 relative, nu and phi
1e-7, branch lengths 1e-6 relative (the shared forest's segment width and
launch shapes change the order of sums over sites)."""
import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.serving import PMMServer, SolveRequest, latency_summary, solve_alone, solve_batch

pytestmark = pytest.mark.gpu


def _requests():
    out = []
    for i, (n, K) in enumerate([(200, 20), (3000, 150), (800, 40), (50, 8), (5000, 300), (1200, 64),
                                (400, 33), (2500, 100), (150, 480), (900, 12)]):
        nu, phi = 0.02 + 0.02 * i, 0.05 + 0.025 * i
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 40, 1.0, nu, phi, 4000 + i)
        spec = sim.problem_spec(tree, Q, codes)
        out.append(SolveRequest(spec, [np.full(tree.n_nodes, 0.05)], 0.5, 0.5, max_iter=60, tol=1e-8, rid=i))
    return out


def test_served_results_equal_solving_alone():
    reqs = _requests()
    alone = {r.rid: solve_alone(r) for r in reqs}
    sizes = sorted(r.bytes for r in reqs)
    budget = sum(sizes[-3:])  # the largest requests cannot all be resident together
    with PMMServer(slots=4, hbm_budget=budget) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=300) for f in futs]
    assert srv.peak_in_use <= budget and srv.peak_active >= 2
    for r in res:
        a = alone[r.rid]
        print("request", r.rid, r.nll, a[0], r.iters, flush=True)
        assert r.nll == a[0] and r.nu == a[2] and r.phi == a[3] and r.iters == a[4]
        for x, y in zip(r.lengths, a[1]):
            np.testing.assert_array_equal(x, y)
    s = latency_summary(res)
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_request_over_budget_rejected():
    r = _requests()[4]
    with PMMServer(slots=1, hbm_budget=r.bytes - 1) as srv:
        with pytest.raises(ValueError):
            srv.submit(r)


def test_concurrent_requests_with_chain_plans():
    """Requests big enough for outside chains by default, served four at a
    time: each worker builds its forest's chain plan while others capture
    their EM steps; results equal solving alone, bit for bit."""
    reqs = []
    for i, (n, K) in enumerate([(7000, 300), (9000, 256), (12000, 192), (6600, 320)]):
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 40, 1.0, 0.1, 0.1, 4100 + i)
        spec = sim.problem_spec(tree, Q, codes)
        reqs.append(SolveRequest(spec, [np.full(tree.n_nodes, 0.05)], 0.5, 0.5, max_iter=20, tol=1e-8, rid=i))
    alone = {r.rid: solve_alone(r) for r in reqs}
    with PMMServer(slots=4) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=300) for f in futs]
    for r in res:
        a = alone[r.rid]
        print("request", r.rid, r.nll, a[0], flush=True)
        assert r.nll == a[0] and r.nu == a[2] and r.phi == a[3] and r.iters == a[4]
        for x, y in zip(r.lengths, a[1]):
            np.testing.assert_array_equal(x, y)


def _close(got, want):
    nll, lengths, nu, phi, iters = got
    assert abs(nll - want[0]) <= 1e-9 * abs(want[0])
    assert abs(nu - want[2]) <= 1e-7 and abs(phi - want[3]) <= 1e-7
    assert iters == want[4]
    for x, y in zip(lengths, want[1]):
        np.testing.assert_allclose(x, y, rtol=1e-6, atol=1e-10)


def test_solve_batch_equals_solving_alone():
    reqs = [r for r in _requests() if r.rid != 4]
    alone = {r.rid: solve_alone(r) for r in reqs}
    got = solve_batch(reqs)
    for r, g in zip(reqs, got):
        print("batched request", r.rid, g[0], alone[r.rid][0], g[4], alone[r.rid][4], flush=True)
        _close(g, alone[r.rid])


def test_server_batches_small_requests():
    reqs = _requests()
    alone = {r.rid: solve_alone(r) for r in reqs}
    cap = 4 * sorted(r.bytes for r in reqs)[-2]   # all but the largest request are small
    with PMMServer(slots=2, batch_bytes=cap, batch_max=8) as srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=300) for f in futs]
    print("batches", srv.batches, flush=True)
    assert max(srv.batches) > 1 and sum(srv.batches) == len(reqs)
    for r in res:
        _close((r.nll, r.lengths, r.nu, r.phi, r.iters), alone[r.rid])


def test_footprint_estimate_bounds_device_use():
    """The admission estimate of a request plus the per-forest allowance the
    server reserves per slot is at least the device memory its forest takes
    through creation and an EM call (free memory before and after, with the
    caching allocator emptied first; a first forest loads the code objects
    and static arrays beforehand)."""
    from laml_amd.forest import Forest
    from laml_amd.serving import FOREST_FIXED_BYTES, estimate_forest_bytes

    dev = torch.cuda.current_device()

    def run(n, K):
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 50, 1.0, 0.1, 0.1, 9)
        spec = sim.problem_spec(tree, Q, codes)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        f0, _ = torch.cuda.mem_get_info(dev)
        F = Forest([spec], device=dev)
        d = torch.tensor(F.lengths_to_level([np.full(tree.n_nodes, 0.05)]), dtype=torch.float64, device=F.device)
        nu = torch.tensor([0.5], dtype=torch.float64, device=F.device)
        phi = nu.clone()
        F.em(d, nu, phi, 5, 1e-8)
        torch.cuda.synchronize()
        f1, _ = torch.cuda.mem_get_info(dev)
        F.close()
        return estimate_forest_bytes([spec]) + FOREST_FIXED_BYTES, f0 - f1

    run(300, 8)
    for n, K in [(2000, 20), (2000, 40), (3000, 65), (5000, 300), (20000, 300)]:
        est, used = run(n, K)
        print("footprint", n, K, est, used, flush=True)
        assert est >= used, (n, K, est, used)
