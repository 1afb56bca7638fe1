"""The benchmark's five workload configurations at their stated sizes, on the
device, against the CPU oracle.

1. 250 leaves x 30 sites (M in [5,50], Dirichlet(1), nu = phi = 0.1): NLL at
   the generating parameters, then 50 EM iterations from the uniform start
   (all branches 0.05, nu = phi = 0.5) against the scalar Python oracle
   (EM_solver.py:461-512): every NLL of the trace to 1e-9 relative, fitted
   branch lengths, nu and phi to 1e-6.
2. 1000 leaves x 100 sites (M in [2,100], Dirichlet(0.5), nu = 0.05,
   phi = 0.2): NLL at the generating parameters (1e-9) and E-step counts
   (1e-9) against the C oracle, then 50 EM iterations against the C oracle's
   E-step / M-step loop: trace to 1e-9, fitted parameters to 1e-6.
3. 768 trees (512 NNI variants of a 2000-leaf tree + 256 trees of 100-5000
   leaves, 200 sites, M in [5,50]) in one forest: every tree's EM result
   (NLL, iteration count, lengths, nu, phi) in the batch equals that tree
   solved in a forest of its own, bit for bit, with the same launch shapes
   (level passes, one 64-site chunk per wavefront); 8 sampled trees against
   the C oracle (1e-9 / 1e-6).
4. 10,000 leaves x 300 sites (M in [5,60], Dirichlet(1), nu = phi = 0.1):
   5 EM iterations against the C oracle; the adjoint gradient against the
   Fisher identity on the C oracle's expected counts (1e-9 of the size of the
   terms each branch's formula sums, since they cancel) and against
   central differences (h = 1e-6) on 16 sampled branches, nu and phi;
   level checkpointing every 16 levels bit-identical to full storage
   (gradients and EM) under the same launch shapes.
5. Serving at the stream's size extremes (50,000 and 200 leaves, 500 and 20
   sites, nu in [0.02,0.2], phi in [0.05,0.3]): served together, each result
   equals the same request solved alone, bit for bit; the smallest against
   the C oracle.
"""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest
from oracle import pmm_ref
from oracle.em_oracle import OracleEM, postorder

pytestmark = pytest.mark.gpu

# generator arguments (n, K, M lo, M hi, Dirichlet alpha, nu, phi, seed)
CONFIG1 = (250, 30, 5, 50, 1.0, 0.1, 0.1, 20261203)
CONFIG2 = (1000, 100, 2, 100, 0.5, 0.05, 0.2, 20261204)
CONFIG3 = dict(base_n=2000, variants=512, others=256, n_lo=100, n_hi=5000, K=200, m_lo=5, m_hi=50, alpha=1.0,
               nu=0.1, phi=0.15, seed=20261208)
CONFIG4 = (10000, 300, 5, 60, 1.0, 0.1, 0.1, 20261206)


def _forest(tree, Q, codes, lengths=None):
    return Forest([sim.problem_spec(tree, Q, codes)], device=torch.cuda.current_device())


def _dev(F, x):
    return torch.tensor(x, dtype=torch.float64, device=F.device)


def _ref(tree, Q, codes):
    return pmm_ref.RefProblem(tree.c0, tree.c1, tree.parent, codes, Q, None, 1.0 / (len(Q) * codes.shape[0]))


def _ref_em_trace(p, d, nu, phi, iters):
    """The C oracle's EM as E-step / M-step calls: NLL before each E-step,
    then the NLL after the last M-step."""
    trace = []
    for _ in range(iters):
        nll, counts, R, Rt = p.estep(d, nu, phi)
        trace.append(nll)
        d, nu, phi = p.mstep(counts, R, Rt, d, nu, phi)
    trace.append(p.nll(d, nu, phi))
    return np.asarray(trace), d, nu, phi


def _device_em(F, n_nodes, iters, tol=0.0):
    d = _dev(F, F.lengths_to_level([np.full(n_nodes, 0.05)]))
    nu = torch.full((F.num_problems,), 0.5, dtype=torch.float64, device=F.device)
    phi = torch.full((F.num_problems,), 0.5, dtype=torch.float64, device=F.device)
    trace = torch.zeros((iters + 1) * F.num_problems, dtype=torch.float64, device=F.device)
    nll, it = F.em(d, nu, phi, iters, tol, trace=trace)
    torch.cuda.synchronize()
    return (trace.cpu().numpy(), F.level_to_tree_lengths(d.cpu().numpy())[0], float(nu[0].item()),
            float(phi[0].item()), float(nll[0].item()))


def test_config1_250x30_truth_nll_and_50_em_iterations_vs_python_oracle():
    tree, Q, codes, truth = sim.make_problem(*CONFIG1)
    F = _forest(tree, Q, codes)
    msa = sim.codes_to_charmtrx(tree, codes)
    # NLL at the generating parameters
    nll = F.nll(_dev(F, F.lengths_to_level([truth["d"]])), _dev(F, [0.1]), _dev(F, [0.1])).item()
    o = OracleEM([tree.newick()], msa, Q, nu=0.1, phi=0.1)
    print("config1 truth nll", nll, o.negative_llh(), flush=True)
    assert math.isclose(nll, o.negative_llh(), rel_tol=1e-9)
    # 50 EM iterations from the uniform start
    trace, d, nu, phi, final = _device_em(F, tree.n_nodes, 50)
    o = OracleEM([tree.newick()], msa, Q, nu=0.5, phi=0.5)
    for v in postorder(o.trees[0]):
        v.edge_length = 0.05
    o.az_partition()
    onll, _, _ = o.em_optimization(fixed_iters=50)
    print("config1 em", final, onll, flush=True)
    np.testing.assert_allclose(-trace, o.llh_trace, rtol=1e-9)
    assert math.isclose(final, onll, rel_tol=1e-9)
    assert abs(nu - o.nu) <= 1e-6 and abs(phi - o.phi) <= 1e-6
    np.testing.assert_allclose(d, [v.edge_length for v in postorder(o.trees[0])], rtol=1e-6, atol=1e-9)
    F.close()


def test_config2_1000x100_vs_c_oracle():
    tree, Q, codes, truth = sim.make_problem(*CONFIG2)
    F = _forest(tree, Q, codes)
    p = _ref(tree, Q, codes)
    d_true = _dev(F, F.lengths_to_level([truth["d"]]))
    nu_t, phi_t = _dev(F, [0.05]), _dev(F, [0.2])
    nll = F.nll(d_true, nu_t, phi_t).item()
    rnll = p.nll(truth["d"], 0.05, 0.2)
    print("config2 truth nll", nll, rnll, flush=True)
    assert math.isclose(nll, rnll, rel_tol=1e-9)
    counts, rt, nll2 = F.estep(d_true, nu_t, phi_t)
    _, rcounts, R, Rt = p.estep(truth["d"], 0.05, 0.2)
    c_post = np.empty((F.num_nodes, 5))
    c_post[F.order] = counts.view(-1, 5).cpu().numpy()
    np.testing.assert_allclose(c_post, rcounts, rtol=1e-9, atol=1e-9 * len(Q))
    assert math.isclose(rt.item(), Rt, rel_tol=1e-9)
    # 50 EM iterations from the uniform start
    trace, d, nu, phi, final = _device_em(F, tree.n_nodes, 50)
    rtrace, rd, rnu, rphi = _ref_em_trace(p, np.full(tree.n_nodes, 0.05), 0.5, 0.5, 50)
    print("config2 em", final, rtrace[-1], nu, rnu, phi, rphi, flush=True)
    np.testing.assert_allclose(trace, rtrace, rtol=1e-9)
    assert abs(nu - rnu) <= 1e-6 and abs(phi - rphi) <= 1e-6
    np.testing.assert_allclose(d, rd, rtol=1e-6, atol=1e-9)
    F.close()


def _fixed_plan(F):
    """Level passes with one 64-site chunk per wavefront on every level: the
    launch shapes of a forest then do not depend on the other problems."""
    F.set_pass_mode(0)
    H = F.num_levels
    F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), H)


def test_config3_768_tree_batch_equals_each_tree_alone():
    problems = sim.scoring_workload(**CONFIG3)
    assert len(problems) == 768
    specs = [sim.problem_spec(t, q, c) for (t, q, c) in problems]
    dev = torch.cuda.current_device()
    iters, tol = 10, 1e-8

    def run(F, probs):
        _fixed_plan(F)
        d = _dev(F, F.lengths_to_level([np.full(t.n_nodes, 0.05) for (t, _, _) in probs]))
        P = len(probs)
        nu = torch.full((P,), 0.5, dtype=torch.float64, device=F.device)
        phi = torch.full((P,), 0.5, dtype=torch.float64, device=F.device)
        nll, it = F.em(d, nu, phi, iters, tol)
        torch.cuda.synchronize()
        return (nll.cpu().numpy().copy(), it.cpu().numpy().copy(), F.level_to_tree_lengths(d.cpu().numpy()),
                nu.cpu().numpy().copy(), phi.cpu().numpy().copy())

    F = Forest(specs, device=dev)
    print("config3 forest", F.num_problems, "problems", F.num_nodes, "nodes", flush=True)
    nll, it, lens, nu, phi = run(F, problems)
    F.close()
    for i, spec in enumerate(specs):
        Fi = Forest([spec], device=dev)
        n1, i1, l1, nu1, phi1 = run(Fi, [problems[i]])
        Fi.close()
        if i % 64 == 0:
            print("config3 tree", i, nll[i], n1[0], it[i], flush=True)
        assert nll[i] == n1[0] and it[i] == i1[0] and nu[i] == nu1[0] and phi[i] == phi1[0], i
        np.testing.assert_array_equal(lens[i], l1[0])
    # sampled trees against the C oracle (the same schedule: EM_solver's
    # convergence test, at most `iters` M-steps)
    for i in list(range(0, 512, 128)) + list(range(512, 768, 64)):
        t, q, c = problems[i]
        rn, rit, rd, rnu, rphi = _ref(t, q, c).em(np.full(t.n_nodes, 0.05), 0.5, 0.5, max_iter=iters, tol=tol)
        print("config3 oracle", i, nll[i], rn, it[i], rit, flush=True)
        assert math.isclose(nll[i], rn, rel_tol=1e-9)
        assert it[i] == rit
        assert abs(nu[i] - rnu) <= 1e-6 and abs(phi[i] - rphi) <= 1e-6
        np.testing.assert_allclose(lens[i], rd, rtol=1e-6, atol=1e-9)


@pytest.fixture(scope="module")
def config4():
    tree, Q, codes, truth = sim.make_problem(*CONFIG4)
    F = _forest(tree, Q, codes)
    yield tree, Q, codes, truth, F
    F.close()


def test_config4_em_vs_c_oracle(config4):
    tree, Q, codes, truth, F = config4
    trace, d, nu, phi, final = _device_em(F, tree.n_nodes, 5)
    rtrace, rd, rnu, rphi = _ref_em_trace(_ref(tree, Q, codes), np.full(tree.n_nodes, 0.05), 0.5, 0.5, 5)
    print("config4 em", trace, rtrace, flush=True)
    np.testing.assert_allclose(trace, rtrace, rtol=1e-9)
    assert abs(nu - rnu) <= 1e-6 and abs(phi - rphi) <= 1e-6
    np.testing.assert_allclose(d, rd, rtol=1e-6, atol=1e-9)


def test_config4_gradient_fisher_identity_and_central_differences(config4):
    tree, Q, codes, truth, F = config4
    rng = np.random.default_rng(4)
    L = truth["d"] * np.exp(0.2 * rng.standard_normal(tree.n_nodes))
    x, y = 0.1, 0.1
    d = _dev(F, F.lengths_to_level([L]))
    nu, phi = _dev(F, [x]), _dev(F, [y])
    nll0 = F.nll(d, nu, phi).item()
    gd, gnu, gphi = F.grad(d, nu, phi)
    gd_post = np.empty(F.num_nodes)
    gd_post[F.order] = gd.cpu().numpy()
    # Fisher identity on the C oracle's expected counts (post-order rows)
    _, S, R, Rt = _ref(tree, Q, codes).estep(L, x, y)
    terms = np.stack([-(1 + x) * S[:, 0], S[:, 1] / np.expm1(L), -x * S[:, 1],
                      (S[:, 2] + S[:, 4]) * x / np.expm1(x * L), -x * S[:, 3]])
    g_d = terms.sum(axis=0)
    # 1e-9 relative to the size of the terms the sum cancels
    err = np.abs(gd_post + g_d)
    scale = np.abs(terms).sum(axis=0)
    assert (err <= 1e-9 * scale).all(), (err / scale).max()
    g_nu = float(np.sum(L * (-S[:, 0] - S[:, 1] - S[:, 3] + (S[:, 2] + S[:, 4]) / np.expm1(x * L))))
    assert math.isclose(gnu.item(), -g_nu, rel_tol=1e-9)
    nobs = float((codes >= 0).sum())
    g_phi = -nobs / (1 - y) + Rt / y
    assert math.isclose(gphi.item(), -g_phi, rel_tol=1e-9)
    # central differences: the NLL's own rounding (~eps * |NLL| per
    # evaluation) divided by 2h bounds the differencing error
    h = 1e-6
    slack = 64 * np.finfo(float).eps * abs(nll0) / h

    def f(dd, nn, pp):
        return F.nll(dd, nn, pp).item()

    for v in rng.choice(F.num_nodes, 16, replace=False):
        dp, dm = d.clone(), d.clone()
        dp[v] += h
        dm[v] -= h
        fd = (f(dp, nu, phi) - f(dm, nu, phi)) / (2 * h)
        assert abs(fd - gd[v].item()) <= 1e-6 * abs(fd) + slack, (v, fd, gd[v].item())
    fd_nu = (f(d, nu + h, phi) - f(d, nu - h, phi)) / (2 * h)
    fd_phi = (f(d, nu, phi + h) - f(d, nu, phi - h)) / (2 * h)
    print("config4 fd", fd_nu, gnu.item(), fd_phi, gphi.item(), "slack", slack, flush=True)
    assert abs(fd_nu - gnu.item()) <= 1e-6 * abs(fd_nu) + slack
    assert abs(fd_phi - gphi.item()) <= 1e-6 * abs(fd_phi) + slack


def test_config4_checkpoint_every_16_levels_is_bit_identical(config4):
    tree, Q, codes, truth, F = config4
    d = _dev(F, F.lengths_to_level([truth["d"]]))
    nu, phi = _dev(F, [0.1]), _dev(F, [0.1])
    # the same launch shapes in both storage modes (checkpointing runs level
    # passes; the default plan's walk and multi-chunk wavefronts would change
    # the order of the sums over sites)
    _fixed_plan(F)
    F.set_checkpoint(0)
    full_rows = F.inside_rows()
    n_full = F.nll(d, nu, phi).item()
    g_full = [t.clone() for t in F.grad(d, nu, phi)]
    d1, nu1, phi1 = d.clone(), nu.clone(), phi.clone()
    F.em(d1, nu1, phi1, 3, 0.0)
    F.set_checkpoint(16)
    ck_rows = F.inside_rows()
    n_ck = F.nll(d, nu, phi).item()
    g_ck = F.grad(d, nu, phi)
    print("config4 checkpoint rows", full_rows, ck_rows, flush=True)
    assert ck_rows < full_rows and n_ck == n_full
    for a, b in zip(g_full, g_ck):
        torch.testing.assert_close(a, b, rtol=0, atol=0)
    d2, nu2, phi2 = d.clone(), nu.clone(), phi.clone()
    F.em(d2, nu2, phi2, 3, 0.0)
    torch.testing.assert_close(d1, d2, rtol=0, atol=0)
    torch.testing.assert_close(nu1, nu2, rtol=0, atol=0)
    torch.testing.assert_close(phi1, phi2, rtol=0, atol=0)
    F.set_checkpoint(0)


def test_config5_serving_size_extremes_equal_alone():
    from laml_amd.serving import PMMServer, SolveRequest, solve_alone

    reqs = []
    for i, (n, K, nu, phi) in enumerate([(50000, 500, 0.2, 0.05), (200, 20, 0.02, 0.3), (50000, 20, 0.02, 0.3),
                                         (200, 500, 0.2, 0.05), (3000, 250, 0.1, 0.15)]):
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 50, 1.0, nu, phi, 20261209 + i)
        reqs.append((tree, Q, codes, SolveRequest(sim.problem_spec(tree, Q, codes), [np.full(tree.n_nodes, 0.05)],
                                                  0.5, 0.5, max_iter=50, tol=1e-8, rid=i)))
    alone = {r.rid: solve_alone(r) for (_, _, _, r) in reqs}
    with PMMServer(slots=4) as srv:
        futs = [srv.submit(r) for (_, _, _, r) in reqs]
        res = [f.result(timeout=600) for f in futs]
    for r in res:
        a = alone[r.rid]
        print("config5 request", r.rid, r.nll, a[0], r.iters, flush=True)
        assert r.nll == a[0] and r.nu == a[2] and r.phi == a[3] and r.iters == a[4]
        for x, y in zip(r.lengths, a[1]):
            np.testing.assert_array_equal(x, y)
    tree, Q, codes, r = reqs[1]
    rn, rit, rd, rnu, rphi = _ref(tree, Q, codes).em(np.full(tree.n_nodes, 0.05), 0.5, 0.5, max_iter=50, tol=1e-8)
    a = alone[1]
    assert math.isclose(a[0], rn, rel_tol=1e-9) and a[4] == rit
    assert abs(a[2] - rnu) <= 1e-6 and abs(a[3] - rphi) <= 1e-6
    np.testing.assert_allclose(a[1][0], rd, rtol=1e-6, atol=1e-9)
