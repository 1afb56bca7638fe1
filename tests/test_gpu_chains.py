"""Heavy-path chains for the outside pass on the GPU: a wavefront walks a node
and its successive heavy children at a site chunk, keeping the outside state
in registers.  With one chunk per wavefront in the level launches the chain
pass does the same operations in the same order, so the counts, dropout
sums, likelihoods, EM trajectories and gradients are the same bits; the
posteriors and quartet EM, which read every node's outside row, recompute
them and agree bit for bit as well.  Cases: Yule trees of several sizes,
ragged site counts (partial chunks), batches of trees with different site
counts, a caterpillar (one chain as long as the tree), two-leaf trees, the
level and hybrid modes with and without the top-level kernel, converged
problems masked in the EM loop."""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest
from oracle.em_oracle import OracleEM

pytestmark = pytest.mark.gpu


def _caterpillar(n, rng):
    """((((l0, l1), l2), l3) ...) in post-order, random branch lengths."""
    c0, c1, par = [], [], []

    def node(a, b):
        c0.append(a)
        c1.append(b)
        par.append(-1)
        v = len(c0) - 1
        if a >= 0:
            par[a] = v
            par[b] = v
        return v

    cur = node(-1, -1)
    for _ in range(n - 1):
        leaf = node(-1, -1)
        cur = node(cur, leaf)
    N = len(c0)
    length = rng.uniform(0.01, 0.2, size=N)
    nl = sum(1 for a in c0 if a < 0)
    return sim.YuleTree(np.array(c0), np.array(c1), np.array(par), length, ["c%d" % i for i in range(nl)])


def _problem(kind, n, K, seed):
    if kind == "yule":
        return sim.make_problem(n, K, 3, 20, 1.0, 0.1, 0.15, seed)
    rng = np.random.Generator(np.random.PCG64(seed))
    t = _caterpillar(n, rng)
    Q = sim.random_priors(K, 3, 20, 1.0, rng)
    codes = sim.simulate_codes(t, Q, 0.1, 0.15, rng)
    return t, Q, codes, None


def _pair(problems, mode, top=None, which=True, tops=(0, 0)):
    """The same forest twice, chains on (both passes, or the flags `which`,
    with chain-region tops `tops` = (inside, outside), 0: up to the split)
    and off, in pass mode `mode` with one chunk per wavefront on every level
    (top: the split level; default all levels in level launches)."""
    specs = [sim.problem_spec(t, q, c) for (t, q, c, _) in problems]
    dev = torch.cuda.current_device()
    out = []
    for on in (which, False):
        F = Forest(specs, device=dev)
        F.set_pass_mode(mode)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), H if top is None else min(top, H))
        F.set_chains(on, *tops)
        out.append(F)
    rows = max(len(q) for (_, q, _, _) in problems) > 32
    want = (3 if which is True else which) if rows else 0
    assert out[0].chains() == want and out[1].chains() == 0
    return out


def _params(F, problems, nu=0.2, phi=0.1, start=None):
    L = [t.length if start is None else np.full(t.n_nodes, start) for (t, _, _, _) in problems]
    d = torch.tensor(F.lengths_to_level(L), dtype=torch.float64, device=F.device)
    P = len(problems)
    return (d, torch.full((P,), nu, dtype=torch.float64, device=F.device),
            torch.full((P,), phi, dtype=torch.float64, device=F.device))


def _estep(F, prm):
    counts, rt, nll = F.estep(*prm)
    return nll.cpu().numpy().copy(), counts.cpu().numpy().copy(), rt.cpu().numpy().copy()


def _close(*Fs):
    for F in Fs:
        F.close()


CASES = [  # (tree kind, leaves, sites, problems)
    ("yule", 300, 150, 1), ("yule", 500, 200, 3), ("yule", 257, 65, 2), ("yule", 1200, 100, 2),
    ("caterpillar", 400, 96, 1), ("yule", 2, 70, 3), ("yule", 150, 20, 1)]


@pytest.mark.parametrize("kind,n,K,P", CASES)
@pytest.mark.parametrize("mode,top,which,tops", [(0, None, 3, (0, 0)), (0, 6, 1, (0, 3)), (2, 6, 3, (4, 0)),
                                                  (2, None, 2, (0, 0)), (0, 7, 2, (3, 5)), (2, 9, 3, (5, 6))])
def test_chain_estep_equals_levels(kind, n, K, P, mode, top, which, tops):
    probs = [_problem(kind, n + 37 * i, K - 9 * i, 1300 + i) for i in range(P)]
    print("case", kind, n, K, P, mode, top, which, tops, flush=True)
    Fa, Fb = _pair(probs, mode, top, which, tops)
    na, ca, ra = _estep(Fa, _params(Fa, probs))
    nb, cb, rb = _estep(Fb, _params(Fb, probs))
    np.testing.assert_array_equal(na, nb)
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(ra, rb)
    t, q, c, _ = probs[0]
    o = OracleEM([t.newick()], sim.codes_to_charmtrx(t, c), q, nu=0.2, phi=0.1)
    assert math.isclose(float(na[0]), o.negative_llh(), rel_tol=1e-9)
    _close(Fa, Fb)


def test_inside_chain_nll_and_site_likelihoods():
    probs = [_problem("yule", 700, 130, 1450), _problem("caterpillar", 300, 70, 1451)]
    Fa, Fb = _pair(probs, 2, 9, 2)
    out = []
    for F in (Fa, Fb):
        d, nu, phi = _params(F, probs)
        nll = F.nll(d, nu, phi, site_llh=True)
        out.append((nll.cpu().numpy().copy(), F.site_llh.cpu().numpy().copy()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])
    _close(Fa, Fb)


def test_chain_em_graph_and_masking():
    # problems of different sizes converge at different steps: the chain
    # kernels skip masked problems like the level kernels
    probs = [_problem("yule", 300 + 250 * i, 120 - 20 * i, 1400 + i) for i in range(4)]
    for mode in (0, 2):
        Fa, Fb = _pair(probs, mode, 8, True, (4, 6))
        out = []
        for F in (Fa, Fb):
            d, nu, phi = _params(F, probs, nu=0.5, phi=0.5, start=0.05)
            nll, it = F.em(d, nu, phi, 60, 1e-7)
            out.append((nll.cpu().numpy().copy(), it.cpu().numpy().copy(), d.cpu().numpy().copy(),
                        nu.cpu().numpy().copy(), phi.cpu().numpy().copy()))
        for a, b in zip(out[0], out[1]):
            np.testing.assert_array_equal(a, b)
        print("mode", mode, "iters", out[0][1], flush=True)
        # a second call replays the captured graph from the same start
        d, nu, phi = _params(Fa, probs, nu=0.5, phi=0.5, start=0.05)
        nll2, _ = Fa.em(d, nu, phi, 60, 1e-7)
        np.testing.assert_array_equal(nll2.cpu().numpy(), out[0][0])
        _close(Fa, Fb)


def test_chain_gradient_equals_levels():
    probs = [_problem("yule", 350, 100, 1500)]
    Fa, Fb = _pair(probs, 2, 7)
    g = []
    for F in (Fa, Fb):
        d, nu, phi = _params(F, probs)
        F.nll(d, nu, phi)
        gd, gnu, gphi = F.grad(d, nu, phi)
        g.append((gd.cpu().numpy().copy(), gnu.item(), gphi.item()))
    np.testing.assert_array_equal(g[0][0], g[1][0])
    assert g[0][1] == g[1][1] and g[0][2] == g[1][2]
    _close(Fa, Fb)


def test_posteriors_after_chain_estep():
    probs = [_problem("yule", 400, 90, 1600), _problem("yule", 250, 70, 1601)]
    Fa, Fb = _pair(probs, 2, 6)
    post = []
    for F in (Fa, Fb):
        prm = _params(F, probs)
        F.estep(*prm)
        p0, pS = F.posteriors(prm[2])
        post.append((np.concatenate(p0), np.concatenate(pS)))
    np.testing.assert_array_equal(post[0][0], post[1][0])
    np.testing.assert_array_equal(post[0][1], post[1][1])
    # the chained pass still runs after the recompute
    assert Fa.chains()
    na, ca, _ = _estep(Fa, _params(Fa, probs))
    nb, cb, _ = _estep(Fb, _params(Fb, probs))
    np.testing.assert_array_equal(ca, cb)
    _close(Fa, Fb)


def test_quartet_scores_with_chains(monkeypatch):
    from laml_amd.quartet import QuartetScorer
    from laml_amd.solvers import DEFAULT_STRATEGY, EM_solver
    from laml_amd.topology_search import Topology_search_batched
    import random

    tree, Q, codes, _ = sim.make_problem(120, 80, 3, 15, 1.0, 0.1, 0.15, 17)
    msa = sim.codes_to_charmtrx(tree, codes)
    random.seed(17)
    s = EM_solver([tree.newick()], {"charMtrx": msa}, {"Q": Q}, {"nu": 0, "phi": 0})
    s.score_tree(strategy=dict(DEFAULT_STRATEGY, fixed_brlen=None))
    res = []
    for on in ("1", "0"):
        monkeypatch.setenv("LAML_CHAINS", on)
        ts = Topology_search_batched(s.get_tree_newick(), EM_solver, data={"charMtrx": msa}, prior={"Q": Q},
                                     params=s.get_params())
        random.seed(1)
        moves = ts._list_moves()[::7]
        scorer = QuartetScorer(ts, DEFAULT_STRATEGY, EM_solver)
        sc = scorer.score(moves)
        res.append(([x for x, _ in sc], [list(l.lengths) for _, l in sc]))
        scorer.close()
    assert res[0] == res[1]


def test_chain_eligibility():
    probs = [_problem("yule", 120, 70, 1700)]
    Fa, Fb = _pair(probs, 0)
    assert Fa.chains() == 3
    Fa.set_chains(1)
    assert Fa.chains() == 1
    with pytest.raises(Exception):
        Fa.set_chains(4)
    with pytest.raises(Exception):
        Fa.set_chains(1, -1, 0)
    Fa.set_chains(True, 3, 4)
    assert Fa.get_chains() == (3, 3, 4)
    Fa.set_pass_mode(1)
    assert not Fa.chains()
    Fa.set_pass_mode(2)
    assert Fa.chains()
    Fa.set_checkpoint(2)
    assert not Fa.chains()
    Fa.set_checkpoint(0)
    Fa.set_pass_mode(0)
    assert Fa.chains()
    Fa.set_fusion(True)
    assert Fa.fusion() and not Fa.chains()
    Fa.set_fusion(False)
    assert Fa.chains()
    _close(Fa, Fb)


def test_default_plan_with_chains_matches_c_oracle_em():
    """A forest big enough that outside chains are on by default (7000
    leaves x 300 sites) runs the same EM as the C oracle (oracle/pmm_ref.c):
    five whole iterations from the benchmark's start."""
    from oracle import pmm_ref

    t, q, c, _ = sim.make_problem(7000, 300, 5, 30, 1.0, 0.1, 0.1, 2024)
    F = Forest([sim.problem_spec(t, q, c)], device=torch.cuda.current_device())
    assert F.chains() == 1
    d = torch.tensor(F.lengths_to_level([np.full(t.n_nodes, 0.05)]), dtype=torch.float64, device=F.device)
    nu = torch.full((1,), 0.5, dtype=torch.float64, device=F.device)
    phi = torch.full((1,), 0.5, dtype=torch.float64, device=F.device)
    nll, _ = F.em(d, nu, phi, 5, 0.0)
    got_nll = float(nll.cpu().numpy()[0])
    got_d = F.level_to_tree_lengths(d.cpu().numpy())[0]
    p = pmm_ref.RefProblem(t.c0, t.c1, t.parent, c, q, None, 1.0 / (len(q) * c.shape[0]))
    rn, _, rd, rnu, rphi = p.em(np.full(t.n_nodes, 0.05), 0.5, 0.5, max_iter=5, tol=0.0, fixed_iters=5)
    print("device", got_nll, "oracle", rn, flush=True)
    assert math.isclose(got_nll, rn, rel_tol=1e-9)
    assert abs(nu.item() - rnu) <= 1e-6 * max(1.0, rnu)
    assert abs(phi.item() - rphi) <= 1e-6 * max(1.0, rphi)
    np.testing.assert_allclose(got_d, rd, rtol=1e-6, atol=1e-9)
    F.close()


def test_ultrametric_em_with_chains_equals_levels():
    """The ultrametric M-step on top of chain passes: the same EM trace and
    parameters, bit for bit, as on level passes."""
    from laml_amd.solvers import EM_solver

    t, q, c, _ = sim.make_problem(150, 80, 3, 15, 1.0, 0.1, 0.1, 1900)
    msa = sim.codes_to_charmtrx(t, c)
    out = []
    for on in (True, False):
        s = EM_solver([t.newick()], {"charMtrx": msa}, {"Q": q}, {"nu": 0.1, "phi": 0.05})
        F = s._forest_for()
        F.set_pass_mode(0)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), H)
        F.set_chains(on, 0, 0)
        assert bool(F.chains()) == on
        d, nu_t, phi_t = s._device_params(F)
        assert s._enable_ultra(F, d.cpu().numpy())
        trace = torch.zeros(6, dtype=torch.float64, device=F.device)
        F.em(d, nu_t, phi_t, 5, 0.0, trace=trace)
        out.append((trace.cpu().numpy().copy(), d.cpu().numpy().copy(), nu_t.cpu().numpy().copy(),
                    phi_t.cpu().numpy().copy()))
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_chains_multi_tree_problem_and_heavy_dropout():
    """Problems of several trees (one shared nu, phi) and data with heavy
    silencing and dropout (many '?' labels): chain passes equal level
    passes bit for bit, E-step and EM."""
    from laml_amd.forest import ProblemSpec

    rng = np.random.Generator(np.random.PCG64(1950))
    q = sim.random_priors(96, 3, 20, 1.0, rng)
    trees = [sim.yule_tree(n, rng) for n in (400, 250, 90)]
    codes = [sim.simulate_codes(t, q, 0.9, 0.6, rng) for t in trees]
    parts = [sim.problem_spec(t, q, c) for t, c in zip(trees, codes)]
    multi = ProblemSpec([p.trees[0] for p in parts], parts[0].sites, sum(p.nobs for p in parts),
                        min(p.phimin for p in parts))
    t4, q4, c4, _ = sim.make_problem(333, 70, 3, 15, 1.0, 1.2, 0.5, 1951)
    specs = [multi, sim.problem_spec(t4, q4, c4)]
    lengths = [t.length for t in trees] + [t4.length]
    dev = torch.cuda.current_device()
    res = []
    for on in (True, False):
        F = Forest(specs, device=dev)
        F.set_pass_mode(2)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), min(7, H))
        F.set_chains(on, 0, 0)
        assert bool(F.chains()) == on
        d = torch.tensor(F.lengths_to_level(lengths), dtype=torch.float64, device=F.device)
        nu = torch.tensor([0.9, 1.2], dtype=torch.float64, device=F.device)
        phi = torch.tensor([0.6, 0.5], dtype=torch.float64, device=F.device)
        counts, rt, nll = F.estep(d, nu, phi)
        e = (counts.cpu().numpy().copy(), rt.cpu().numpy().copy(), nll.cpu().numpy().copy())
        nll2, _ = F.em(d, nu, phi, 10, 0.0)
        res.append(e + (nll2.cpu().numpy().copy(), d.cpu().numpy().copy(), nu.cpu().numpy().copy(),
                        phi.cpu().numpy().copy()))
        F.close()
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("nprob", [4, 9])
def test_em_batched_problems_equal_each_alone(nprob):
    """Problems of one forest converge at different EM steps; every problem
    gets the results of a forest of its own, bit for bit (level passes, one
    chunk per wavefront).  Four problems use the multi-block nu solve (whose
    launches after a problem converges or finishes its EM exit early), nine
    the one-block solve, which sums in the multi-block solve's order, while
    each problem alone uses the multi-block solve."""
    probs = [_problem("yule", 200 + 150 * (i % 4), 90 - 15 * (i % 4) - i, 1700 + i) for i in range(nprob)]
    dev = torch.cuda.current_device()

    def run(ps):
        F = Forest([sim.problem_spec(t, q, c) for (t, q, c, _) in ps], device=dev)
        F.set_pass_mode(2)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), H)
        d, nu, phi = _params(F, ps, nu=0.5, phi=0.5, start=0.05)
        nll, it = F.em(d, nu, phi, 60, 1e-7)
        out = (nll.cpu().numpy().copy(), it.cpu().numpy().copy(), F.level_to_tree_lengths(d.cpu().numpy()),
               nu.cpu().numpy().copy(), phi.cpu().numpy().copy())
        F.close()
        return out

    nll, it, lens, nu, phi = run(probs)
    assert len(set(it.tolist())) > 1, it
    for i, p in enumerate(probs):
        n1, i1, l1, nu1, phi1 = run([p])
        assert nll[i] == n1[0] and it[i] == i1[0] and nu[i] == nu1[0] and phi[i] == phi1[0], i
        np.testing.assert_array_equal(lens[i], l1[0])


@pytest.mark.parametrize("cap", [2, 3, 8, 40])
def test_outside_packs_equal_one_chain_per_wavefront(cap, monkeypatch):
    """Outside packs (a wavefront runs several chains of one tree back to
    back, each later head's outside state staged during the step before it)
    at several pack sizes give the bits of one chain per wavefront and of
    level passes: E-step counts, dropout sums, likelihoods and an EM run of
    8 steps, on four trees of three problems with ragged site counts."""
    from laml_amd.forest import ProblemSpec

    rng = np.random.Generator(np.random.PCG64(2207))
    q = sim.random_priors(130, 3, 20, 1.0, rng)
    trees = [sim.yule_tree(n, rng) for n in (900, 380)]
    codes = [sim.simulate_codes(t, q, 0.2, 0.3, rng) for t in trees]
    parts = [sim.problem_spec(t, q, c) for t, c in zip(trees, codes)]
    multi = ProblemSpec([p.trees[0] for p in parts], parts[0].sites, sum(p.nobs for p in parts),
                        min(p.phimin for p in parts))
    probs = [_problem("yule", 700, 75, 2208), _problem("caterpillar", 300, 140, 2209)]
    specs = [multi] + [sim.problem_spec(t, q2, c) for (t, q2, c, _) in probs]
    lengths = [t.length for t in trees] + [t.length for (t, _, _, _) in probs]
    dev = torch.cuda.current_device()
    res = []
    # pack every round up to the cap (small forests keep one chain per
    # wavefront by default)
    monkeypatch.setenv("LAML_PACK_WAVES", "1")
    for pack, on in ((cap, True), (1, True), (1, False)):
        monkeypatch.setenv("LAML_CHAIN_PACK", str(pack))
        print("outside packs", pack, on, flush=True)
        F = Forest(specs, device=dev)
        F.set_pass_mode(0)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), H)
        F.set_chains(1 if on else False, 0, 0)
        assert F.chains() == (1 if on else 0)
        d = torch.tensor(F.lengths_to_level(lengths), dtype=torch.float64, device=F.device)
        nu = torch.tensor([0.2, 0.1, 0.3], dtype=torch.float64, device=F.device)
        phi = torch.tensor([0.3, 0.15, 0.1], dtype=torch.float64, device=F.device)
        counts, rt, nll = F.estep(d, nu, phi)
        e = (counts.cpu().numpy().copy(), rt.cpu().numpy().copy(), nll.cpu().numpy().copy())
        nll2, _ = F.em(d, nu, phi, 8, 0.0)
        res.append(e + (nll2.cpu().numpy().copy(), d.cpu().numpy().copy(), nu.cpu().numpy().copy(),
                        phi.cpu().numpy().copy()))
        F.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            np.testing.assert_array_equal(a, b)


def test_chain_prior_weight_rows_with_unnormalised_priors(monkeypatch):
    """Per-site prior masses other than 1 (the fastEM model's observed-state
    alphabets): the prior-weight rows hold the mass for labels that name no
    edit state, and chain passes (packed and not) still equal level passes
    bit for bit, E-step and EM."""
    from laml_amd.forest import SiteTable

    rng = np.random.Generator(np.random.PCG64(2301))
    tree, Q, codes, _ = sim.make_problem(1500, 150, 4, 25, 1.0, 0.1, 0.2, 2302)
    scale = rng.uniform(0.3, 0.9, size=len(Q))
    Qs = [{s: p * scale[k] for s, p in Q[k].items()} for k in range(len(Q))]
    sites = SiteTable(Qs, normalized=False)
    assert np.abs(sites.qsum - 1.0).min() > 0.05
    spec = sim.problem_spec(tree, Qs, codes, sites=sites)
    dev = torch.cuda.current_device()
    monkeypatch.setenv("LAML_PACK_WAVES", "1")
    res = []
    for on, pack in ((1, 8), (1, 1), (False, 1)):
        monkeypatch.setenv("LAML_CHAIN_PACK", str(pack))
        print("unnormalised priors, chains", on, "pack", pack, flush=True)
        F = Forest([spec], device=dev)
        F.set_pass_mode(2)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), min(9, H))
        F.set_chains(on, 0, 0)
        assert F.chains() == (1 if on else 0)
        d = torch.tensor(F.lengths_to_level([tree.length]), dtype=torch.float64, device=F.device)
        nu = torch.full((1,), 0.1, dtype=torch.float64, device=F.device)
        phi = torch.full((1,), 0.2, dtype=torch.float64, device=F.device)
        counts, rt, nll = F.estep(d, nu, phi)
        e = (counts.cpu().numpy().copy(), rt.cpu().numpy().copy(), nll.cpu().numpy().copy())
        nll2, _ = F.em(d, nu, phi, 6, 0.0)
        res.append(e + (nll2.cpu().numpy().copy(), d.cpu().numpy().copy(), nu.cpu().numpy().copy(),
                        phi.cpu().numpy().copy()))
        F.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            np.testing.assert_array_equal(a, b)


def test_benchmark_batch_chain_passes_equal_levels():
    """The benchmark's batch (16 problems, each a 10,000-leaf Yule tree x 300
    sites) with the default pack settings, so that the big chain rounds run
    packed and the small ones a chain per wavefront: the E-step equals level
    passes with one chunk per wavefront bit for bit."""
    probs = [sim.make_problem(10000, 300, 5, 60, 1.0, 0.1, 0.1, 20261206 + 1000 * b) for b in range(16)]
    specs = [sim.problem_spec(t, q, c) for (t, q, c, _) in probs]
    lengths = [np.full(t.n_nodes, 0.05) for (t, _, _, _) in probs]
    dev = torch.cuda.current_device()
    res = []
    for on in (1, False):
        print("benchmark batch, chains", on, flush=True)
        F = Forest(specs, device=dev)
        F.set_pass_mode(2)
        H = F.num_levels
        F.set_shapes(np.ones(H, dtype=np.int32), np.ones(H, dtype=np.int32), 18)
        F.set_chains(on, 0, 0)
        assert F.chains() == (1 if on else 0)
        d = torch.tensor(F.lengths_to_level(lengths), dtype=torch.float64, device=F.device)
        nu = torch.full((16,), 0.5, dtype=torch.float64, device=F.device)
        phi = torch.full((16,), 0.5, dtype=torch.float64, device=F.device)
        counts, rt, nll = F.estep(d, nu, phi)
        res.append((counts.cpu().numpy().copy(), rt.cpu().numpy().copy(), nll.cpu().numpy().copy()))
        F.close()
    for a, b in zip(*res):
        np.testing.assert_array_equal(a, b)
