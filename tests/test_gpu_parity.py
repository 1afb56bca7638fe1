"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Tolerances: NLL 1e-9 relative (the stated bar; observed agreement is ~1e-13),
per-branch expected counts 1e-9 relative + 1e-12 absolute, fitted parameters
1e-6 absolute after the same EM schedule, az-partition labels bit-exact.
"""
import math

import numpy as np
import pytest
import torch

from helpers import kat_inputs, load_example1, load_kats, load_test1, load_test4
from laml_amd import sim
from laml_amd.solvers import EM_solver
from oracle.em_oracle import OracleEM, postorder

pytestmark = pytest.mark.gpu

KATS = load_kats()
NLL_RTOL = 1e-9


def _pair(T, msa, Q, nu, phi):
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"nu": nu, "phi": phi})
    return o, s


def _synthetic(n, K, mlo, mhi, alpha, nu, phi, seed, start=None):
    tree, Q, codes, _ = sim.make_problem(n, K, mlo, mhi, alpha, nu, phi, seed)
    T = tree.newick()
    msa = sim.codes_to_charmtrx(tree, codes)
    return T, msa, Q


@pytest.mark.parametrize("case", KATS["nllh_kats"], ids=lambda c: c["name"])
def test_kat_nllh(case):
    T, msa, Q, nu, phi = kat_inputs(case)
    o, s = _pair(T, msa, Q, nu, phi)
    want = o.negative_llh()
    got = s.negative_llh()
    print(case["name"], got, want, flush=True)
    assert math.isclose(got, want, rel_tol=1e-12, abs_tol=1e-13)
    assert abs(got - case["nllh"]) < 2e-9


DATASETS = {
    "test1": lambda: load_test1() + (0.15877477685098262, 0.05231954386883335),
    "test4": lambda: load_test4() + (0.1, 0.05),
    "example1": lambda: load_example1() + (0.1, 0.1),
    "syn250x30": lambda: _synthetic(250, 30, 5, 50, 1.0, 0.1, 0.1, 20261203) + (0.1, 0.1),
    "syn300x40_skew": lambda: _synthetic(300, 40, 2, 100, 0.5, 0.05, 0.2, 7) + (0.05, 0.2),
    "syn200x150": lambda: _synthetic(200, 150, 5, 50, 1.0, 0.1, 0.1, 31) + (0.1, 0.1),
}


@pytest.fixture(scope="module", params=sorted(DATASETS))
def dataset(request):
    T, msa, Q, nu, phi = DATASETS[request.param]()
    o, s = _pair(T, msa, Q, nu, phi)
    o.az_partition()
    return request.param, T, msa, Q, nu, phi, o, s


def _post_nodes(s):
    return list(s.trees[0].traverse_postorder())


def test_labels_bit_exact(dataset):
    name, T, msa, Q, nu, phi, o, s = dataset
    print("labels", name, flush=True)
    got = s.node_labels()[0]
    for vo, vs in zip(postorder(o.trees[0]), _post_nodes(s)):
        assert got[vs] == vo.a["alpha"], (name, vo.label)


def test_nllh_and_site_llh(dataset):
    name, T, msa, Q, nu, phi, o, s = dataset
    print("nllh", name, flush=True)
    want = o.negative_llh()
    got = s.negative_llh()
    assert math.isclose(got, want, rel_tol=NLL_RTOL)
    site = s.site_llh()[0]
    ws = o.site_llh()
    np.testing.assert_allclose(site, ws, rtol=1e-10, atol=1e-10)


def test_estep_counts(dataset):
    name, T, msa, Q, nu, phi, o, s = dataset
    print("estep", name, flush=True)
    o.estep()
    nll, rt = s.Estep()
    R, Rt = o.phi_terms()
    assert math.isclose(rt, Rt, rel_tol=1e-9, abs_tol=1e-10)
    for vo, vs in zip(postorder(o.trees[0]), _post_nodes(s)):
        want = [sum(vo.a["S"][j]) for j in range(5)]
        np.testing.assert_allclose(vs.S_sum, want, rtol=1e-9, atol=1e-9 * o.numsites, err_msg=str(vo.label))


def test_mstep_one_step(dataset):
    name, T, msa, Q, nu, phi, o, s = dataset
    print("mstep", name, flush=True)
    o.estep()
    o.mstep()
    s.Estep()
    s.Mstep()
    assert math.isclose(s.params.phi, o.phi, rel_tol=1e-9, abs_tol=1e-12)
    assert math.isclose(s.params.nu, o.nu, rel_tol=1e-8, abs_tol=1e-12)
    for vo, vs in zip(postorder(o.trees[0]), _post_nodes(s)):
        assert math.isclose(vs.edge_length, vo.edge_length, rel_tol=1e-8, abs_tol=1e-12), vo.label


def _uniform_start(solver, oracle, d0=0.05, nu0=0.5, phi0=0.5):
    for v in solver.trees[0].traverse_postorder():
        v.edge_length = d0
    for v in postorder(oracle.trees[0]):
        v.edge_length = d0
    solver.params.nu = oracle.nu = nu0
    solver.params.phi = oracle.phi = phi0


@pytest.mark.parametrize("name", ["example1", "syn250x30"])
def test_em_fixed_schedule(name):
    T, msa, Q, nu, phi = DATASETS[name]()
    o, s = _pair(T, msa, Q, nu, phi)
    _uniform_start(s, o)
    iters = 12
    o.az_partition()
    onll, _, _ = o.em_optimization(fixed_iters=iters)
    F = s._forest_for()
    d, nu_t, phi_t = s._device_params(F)
    trace = torch.zeros((iters + 1) * 1, dtype=torch.float64, device=F.device)
    nll, it = F.em(d, nu_t, phi_t, iters, 0.0, trace=trace)
    s._store_params(F, d, nu_t, phi_t)
    print("em", name, nll.item(), onll, flush=True)
    assert int(it.item()) == iters + 1
    np.testing.assert_allclose((-trace).cpu().numpy(), o.llh_trace, rtol=NLL_RTOL)
    assert math.isclose(nll.item(), onll, rel_tol=NLL_RTOL)
    assert abs(s.params.nu - o.nu) < 1e-6
    assert abs(s.params.phi - o.phi) < 1e-6
    for vo, vs in zip(postorder(o.trees[0]), _post_nodes(s)):
        assert abs(vs.edge_length - vo.edge_length) < 1e-6


def test_em_reference_kat_test47():
    em = KATS["em_kat"]
    T, msa, Q = load_test4()
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"phi": 0, "nu": 0})
    nllh, status = s.optimize(initials=1, random_seeds=em["random_seed"], verbose=-1, ultra_constr=False)
    print("test47", nllh, s.params.phi, s.params.nu, flush=True)
    assert abs(nllh - em["nllh"]) / em["nllh"] < em["nllh_rel_tol"]
    assert round(abs(s.params.phi - em["phi"]), em["param_places"]) == 0
    assert round(abs(s.params.nu - em["nu"]), em["param_places"]) == 0
    # and the same EM run as the oracle, iteration for iteration
    o = OracleEM([T], msa, Q, 0, 0)
    onll, _, onu, ophi, _, _ = o.optimize(initials=1, random_seeds=em["random_seed"])
    assert math.isclose(nllh, onll, rel_tol=NLL_RTOL)
    assert s.em_iterations == o.em_iters


def test_two_trees_one_problem():
    T1, msa, Q, nu, phi = DATASETS["test4"]()
    tree2 = "((((64:0.2,65:0.1):0.3,(70:0.2,71:0.2):0.1):0.2,(124:0.1,125:0.3):0.4):0.1,(100:0.2,101:0.2):0.3):0.05;"
    sub = {c: msa[c] for c in ["64", "65", "70", "71", "124", "125", "100", "101"]}
    msa2 = dict(msa)
    o = OracleEM([T1, tree2], msa2, Q, nu=nu, phi=phi)
    s = EM_solver([T1, tree2], {"charMtrx": msa2}, {"Q": Q}, {"nu": nu, "phi": phi})
    assert math.isclose(s.negative_llh(), o.negative_llh(), rel_tol=NLL_RTOL)
    o.estep()
    o.mstep()
    s.Estep()
    s.Mstep()
    assert math.isclose(s.params.nu, o.nu, rel_tol=1e-8)
    assert math.isclose(s.params.phi, o.phi, rel_tol=1e-9, abs_tol=1e-12)
    del sub


def test_batched_forest_equals_single():
    from laml_amd.forest import Forest

    specs = []
    singles = []
    for i, (n, K) in enumerate([(40, 7), (130, 33), (75, 70), (300, 16)]):
        tree, Q, codes, _ = sim.make_problem(n, K, 3, 30, 1.0, 0.1, 0.1, 100 + i)
        specs.append(sim.problem_spec(tree, Q, codes))
        T = tree.newick()
        o = OracleEM([T], sim.codes_to_charmtrx(tree, codes), Q, nu=0.1, phi=0.1)
        singles.append(o.negative_llh())
    F = Forest(specs, device=torch.cuda.current_device())
    d = torch.tensor(F.lengths_to_level([t.length for p in specs for t in p.trees]), dtype=torch.float64,
                     device=F.device)
    nu = torch.full((len(specs),), 0.1, dtype=torch.float64, device=F.device)
    phi = torch.full((len(specs),), 0.1, dtype=torch.float64, device=F.device)
    got = F.nll(d, nu, phi).cpu().numpy()
    np.testing.assert_allclose(got, singles, rtol=NLL_RTOL)


@pytest.mark.parametrize("n,K,seed", [(10000, 12, 20261206), (3000, 24, 11)])
def test_deep_tree_nll_and_counts(n, K, seed):
    """Deep Yule trees (height ~30): exponent ranges far beyond the small
    fixtures.  Likelihood and one E-step against the oracle."""
    tree, Q, codes, _ = sim.make_problem(n, K, 5, 60, 1.0, 0.1, 0.1, seed)
    T = tree.newick()
    msa = sim.codes_to_charmtrx(tree, codes)
    o = OracleEM([T], msa, Q, nu=0.1, phi=0.1)
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"nu": 0.1, "phi": 0.1})
    want = o.negative_llh()
    got = s.negative_llh()
    print("deep", n, K, got, want, flush=True)
    assert math.isclose(got, want, rel_tol=NLL_RTOL)
    o.estep()
    s.Estep()
    worst = 0.0
    for vo, vs in zip(postorder(o.trees[0]), _post_nodes(s)):
        want_c = np.array([sum(vo.a["S"][j]) for j in range(5)])
        worst = max(worst, float(np.max(np.abs(np.array(vs.S_sum) - want_c))))
    print("deep counts max abs diff", worst, flush=True)
    assert worst < 1e-8 * K


def test_deep_tree_em_from_uniform_start():
    """The benchmark's start (all branches 0.05, nu = phi = 0.5) on a deep
    10k-leaf tree with few sites: EM trajectory against the oracle."""
    tree, Q, codes, _ = sim.make_problem(10000, 8, 5, 60, 1.0, 0.1, 0.1, 20261206)
    T = tree.newick()
    msa = sim.codes_to_charmtrx(tree, codes)
    o, s = _pair(T, msa, Q, 0.5, 0.5)
    _uniform_start(s, o)
    iters = 4
    o.az_partition()
    onll, _, _ = o.em_optimization(fixed_iters=iters)
    F = s._forest_for()
    d, nu_t, phi_t = s._device_params(F)
    trace = torch.zeros(iters + 1, dtype=torch.float64, device=F.device)
    nll, it = F.em(d, nu_t, phi_t, iters, 0.0, trace=trace)
    print("deep em", (-trace).cpu().numpy(), o.llh_trace, flush=True)
    np.testing.assert_allclose((-trace).cpu().numpy(), o.llh_trace, rtol=NLL_RTOL)
    assert abs(nu_t.item() - o.nu) < 1e-6 and abs(phi_t.item() - o.phi) < 1e-6


def _ultra_pair(T, msa, Q, nu, phi, fixed_labels=()):
    o, s = _pair(T, msa, Q, nu, phi)
    o.az_partition()
    return o, s


@pytest.mark.parametrize("name", ["test4", "syn250x30"])
def test_ultrametric_mstep_matches_slsqp(name):
    """One constrained M-step: device barrier-Newton vs scipy SLSQP on the
    restated reference constraint matrix; the device result satisfies the
    equality system to rounding."""
    T, msa, Q, nu, phi = DATASETS[name]()
    o, s = _pair(T, msa, Q, nu, phi)
    o.az_partition()
    o.estep()
    M, b = o.ultrametric_constr()
    o.mstep(ultra=(M, b))
    s.Estep()
    ok, status = s.Mstep(ultra_constr_cache=s.ultrametric_constr())
    assert ok and status == "optimal"
    assert s._forest.ultra_status()[0] == 0
    dg = np.array([v.edge_length for v in _post_nodes(s)])
    do = np.array([v.edge_length for v in postorder(o.trees[0])])
    Ms, bs = s.ultrametric_constr()
    print("ultra", name, np.abs(dg - do).max(), np.abs(np.array(Ms) @ dg - np.array(bs)).max(), flush=True)
    assert np.abs(np.array(Ms) @ dg - np.array(bs)).max() < 1e-10
    assert (dg >= 0.005 - 1e-12).all() and (dg <= 10 + 1e-12).all()
    np.testing.assert_allclose(dg, do, atol=1e-6)
    assert abs(s.params.nu - o.nu) < 1e-6


def test_ultrametric_em_matches_oracle():
    T, msa, Q, nu, phi = DATASETS["test4"]()
    o, s = _pair(T, msa, Q, 0.1, 0.05)
    iters = 5
    o.az_partition()
    onll, _, _ = o.em_optimization(fixed_iters=iters, ultra_constr=True)
    F = s._forest_for()
    d, nu_t, phi_t = s._device_params(F)
    assert s._enable_ultra(F, d.cpu().numpy())
    trace = torch.zeros(iters + 1, dtype=torch.float64, device=F.device)
    nll, it = F.em(d, nu_t, phi_t, iters, 0.0, trace=trace)
    print("ultra em", (-trace).cpu().numpy(), o.llh_trace, flush=True)
    np.testing.assert_allclose((-trace).cpu().numpy(), o.llh_trace, rtol=1e-8)
    s._store_params(F, d, nu_t, phi_t)
    dg = np.array([v.edge_length for v in _post_nodes(s)])
    do = np.array([v.edge_length for v in postorder(o.trees[0])])
    np.testing.assert_allclose(dg, do, atol=1e-6)


def test_ultrametric_optimize_drop_in():
    """optimize(ultra_constr=True) through the drop-in surface: the fitted tree
    is ultrametric (all root-to-leaf paths equal) and its NLL is no better than
    the unconstrained optimum."""
    T, msa, Q = load_test4()
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"phi": 0, "nu": 0})
    nllh_u, status = s.optimize(initials=1, random_seeds=1221, verbose=-1, ultra_constr=True)
    assert status == "optimal"
    tr = s.trees[0]
    dist = {}
    for v in tr.traverse_preorder():
        dist[v] = (v.edge_length or 0.0) + (dist[v.parent] if v.parent is not None else 0.0)
    depths = [dist[v] for v in tr.traverse_leaves()]
    print("ultra optimize", nllh_u, min(depths), max(depths), flush=True)
    assert max(depths) - min(depths) < 1e-8
    assert nllh_u >= 1007.009158767873 * (1 - 1e-6)


def test_launch_shapes_agree():
    """Every launch shape of the level passes (chunks per wavefront, fused
    top levels on or off) gives the same likelihood and counts up to the
    order of the sums over sites, and the oracle's likelihood."""
    from laml_amd.forest import Forest

    tree, Q, codes, _ = sim.make_problem(600, 200, 5, 50, 1.0, 0.1, 0.1, 77)
    o = OracleEM([tree.newick()], sim.codes_to_charmtrx(tree, codes), Q, nu=0.1, phi=0.1)
    want = o.negative_llh()
    F = Forest([sim.problem_spec(tree, Q, codes)], device=torch.cuda.current_device())
    d = torch.tensor(F.lengths_to_level([tree.length]), dtype=torch.float64, device=F.device)
    nu = torch.full((1,), 0.1, dtype=torch.float64, device=F.device)
    phi = torch.full((1,), 0.1, dtype=torch.float64, device=F.device)
    a0, b0, top0 = F.get_shapes()
    H = F.num_levels
    assert a0[1:].min() >= 1  # 200 sites: row kernels
    results = []
    for cpw, top in [(None, None), (1, None), (4, None), (1, H), (4, 2)]:
        if cpw is None:
            F.set_shapes(a0, b0, top0)
        else:
            F.set_shapes(np.full(H, cpw), np.full(H, cpw), top if top is not None else top0)
        counts, rt, nll = F.estep(d, nu, phi)
        results.append((nll.item(), counts.cpu().numpy().copy(), rt.item()))
        print("shape", cpw, top, nll.item(), want, flush=True)
    for nll, counts, rt in results:
        assert math.isclose(nll, want, rel_tol=NLL_RTOL)
        assert math.isclose(nll, results[0][0], rel_tol=1e-13)
        np.testing.assert_allclose(counts, results[0][1], rtol=1e-11, atol=1e-12)
        assert math.isclose(rt, results[0][2], rel_tol=1e-12, abs_tol=1e-12)
    with pytest.raises(Exception):
        F.set_shapes(np.full(H, 99), None, None)
