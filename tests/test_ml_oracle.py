"""The ML_solver restatement (oracle/ml_oracle.py) against the reference's ML
known-answer tests, its legacy polytomy case, brute-force enumeration and the
dense oracle; the host side of laml_amd's ML_solver (polytomy resolution for
the device, constraint matrix, bounds).  CPU only."""
import json
import math
import random

import numpy as np
import pytest

from helpers import golden_path
from laml_amd.forest import resolved_tree_arrays
from laml_amd.solvers import EPS, ML_solver
from laml_amd.tree import read_tree_newick
from oracle import ml_oracle
from oracle.brute import brute_nllh
from oracle.dense import dense_estep
from oracle.em_oracle import read_charmtrx, read_priors_csv

with open(golden_path("ml_kats.json")) as _f:
    ML_KATS = json.load(_f)


def _kat(case):
    Q = [{int(k): v for k, v in q.items()} for q in case["Q"]]
    return case["tree"], case["charMtrx"], Q, case["nu"], case["phi"]


def load_poly_case():
    c = ML_KATS["ml_poly_nllh"]
    T = open(golden_path(c["tree_file"])).read().strip()
    msa = read_charmtrx(golden_path(c["charMtrx_file"]))
    K = len(next(iter(msa.values())))
    Q = read_priors_csv(golden_path(c["prior_file"]), K)
    return T, msa, Q, c["nu"], c["phi"], c["nllh"]


def random_multifurcating(rng, n_leaves, K, m, max_arity=4, p_missing=0.15, p_sil=0.1):
    """Random tree whose internal nodes have 2..max_arity children, random
    priors and a random character matrix with '?' and -1 entries."""
    subs = ["L%d:%.4f" % (i, rng.uniform(0.02, 1.2)) for i in range(n_leaves)]
    ii = 0
    while len(subs) > 1:
        k = min(len(subs), rng.randint(2, max_arity))
        pick = sorted(rng.sample(range(len(subs)), k), reverse=True)
        parts = [subs[i] for i in pick]
        for i in pick:
            subs.pop(i)
        subs.append("(%s)I%d:%.4f" % (",".join(parts), ii, rng.uniform(0.02, 1.2)))
        ii += 1
    T = subs[0] + ";"
    Q = []
    for _ in range(K):
        w = [rng.random() + 0.05 for _ in range(m)]
        Q.append({a + 1: w[a] for a in range(m)})
    msa = {}
    for i in range(n_leaves):
        row = []
        for _ in range(K):
            r = rng.random()
            if r < p_missing:
                row.append("?")
            elif r < p_missing + p_sil:
                row.append(-1)
            else:
                row.append(rng.randint(0, m))
        msa["L%d" % i] = row
    return T, msa, Q


def _as_ml_missing(msa):
    return {c: ["?" if x == -1 else x for x in row] for c, row in msa.items()}


@pytest.mark.parametrize("case", ML_KATS["ml_nllh"], ids=lambda c: c["name"])
def test_ml_oracle_reference_constants(case):
    T, msa, Q, nu, phi = _kat(case)
    assert abs(ml_oracle.negative_llh(T, msa, Q, nu, phi) - case["nllh"]) < 2e-9


def test_ml_oracle_polytomy_case():
    T, msa, Q, nu, phi, want = load_poly_case()
    assert any(len(v.children) > 2 for v in read_tree_newick(T).traverse_postorder())
    assert math.isclose(ml_oracle.negative_llh(T, msa, Q, nu, phi), want, rel_tol=1e-12)


@pytest.mark.parametrize("seed", range(10))
def test_ml_oracle_matches_brute_force_on_polytomies(seed):
    rng = random.Random(100 + seed)
    T, msa, Q = random_multifurcating(rng, 4 + seed % 2, 3, 2, max_arity=3 + seed % 2)
    nu = [0.05, 0.2, 0.7][seed % 3]
    phi = [0.0, 0.15, 0.4][seed % 3]
    want = brute_nllh(T, _as_ml_missing(msa), Q, nu, phi)
    assert math.isclose(ml_oracle.negative_llh(T, msa, Q, nu, phi), want, rel_tol=1e-10, abs_tol=1e-12)


@pytest.mark.parametrize("seed", range(6))
def test_ml_oracle_matches_dense_on_polytomies(seed):
    rng = random.Random(200 + seed)
    T, msa, Q = random_multifurcating(rng, 25, 6, 4, max_arity=6)
    nu, phi = 0.1 + 0.1 * seed, 0.05 * seed
    Qn = ml_oracle.normalise_Q(Q)
    nll, _, _ = dense_estep(T, msa, Qn, nu, phi, missing=("?", -1), silenced=())
    assert math.isclose(ml_oracle.negative_llh(T, msa, Q, nu, phi), nll, rel_tol=1e-10)


def _resolved_newick(ta, leaf_labels):
    leaf_of = {int(i): leaf_labels[j] for j, i in enumerate(ta.leaf_post)}

    def rec(i):
        if ta.c0[i] < 0:
            s = leaf_of[i]
        else:
            s = "(%s,%s)" % (rec(int(ta.c0[i])), rec(int(ta.c1[i])))
        return "%s:%r" % (s, float(ta.length[i]))

    return rec(ta.n_nodes - 1) + ";"


@pytest.mark.parametrize("seed", range(6))
def test_zero_length_resolution_keeps_the_likelihood(seed):
    rng = random.Random(300 + seed)
    T, msa, Q = random_multifurcating(rng, 30, 5, 3, max_arity=7)
    tree = read_tree_newick(T)
    tree.suppress_unifurcations()
    leaves = [v.label for v in tree.traverse_postorder() if v.is_leaf()]
    ta, post_index = resolved_tree_arrays(tree, lambda v: np.zeros(5, np.int16), lambda v: True)
    # binary, children before parents, parents consistent
    inner = np.nonzero(ta.c0 >= 0)[0]
    assert np.all(ta.c1[inner] >= 0)
    assert np.all(ta.c0[inner] < inner) and np.all(ta.c1[inner] < inner)
    assert np.all(ta.parent[ta.c0[inner]] == inner) and np.all(ta.parent[ta.c1[inner]] == inner)
    assert ta.parent[-1] == -1 and np.sum(ta.parent < 0) == 1
    # every original node mapped once; the added nodes are fixed zero-length
    nodes = list(tree.traverse_postorder())
    assert len(set(post_index.tolist())) == len(nodes)
    added = np.setdiff1d(np.arange(ta.n_nodes), post_index)
    n_added = sum(len(v.children) - 2 for v in nodes if len(v.children) > 2)
    assert len(added) == n_added
    assert np.all(ta.length[added] == 0.0) and np.all(ta.free[added] == 0) and np.all(ta.poly[added])
    for v, i in zip(nodes, post_index):
        assert ta.length[i] == (v.edge_length or 0.0)
    # balanced resolution: a node of arity m is at most ceil(log2 m) levels deep
    hgt, h0 = {}, {}
    for v in nodes:
        hgt[id(v)] = 0 if v.is_leaf() else max(hgt[id(c)] for c in v.children) + math.ceil(math.log2(len(v.children)))
        h0[id(v)] = 0 if v.is_leaf() else max(h0[id(c)] for c in v.children) + 1
    assert h0[id(nodes[-1])] <= ta.height.max() <= hgt[id(nodes[-1])]
    nu, phi = 0.3, 0.1
    want = ml_oracle.negative_llh(T, msa, Q, nu, phi)
    got = ml_oracle.negative_llh(_resolved_newick(ta, leaves), msa, Q, nu, phi)
    assert math.isclose(got, want, rel_tol=1e-12)


def test_ml_solver_constraint_matrix_matches_oracle():
    rng = random.Random(7)
    T, msa, Q = random_multifurcating(rng, 12, 3, 2, max_arity=2)
    T2, _, _ = random_multifurcating(random.Random(8), 12, 3, 2, max_arity=2)
    s = ML_solver([T, T2], {"charMtrx": msa}, {"Q": Q}, {"nu": 0.1, "phi": 0.1})
    M = np.asarray(s.ultrametric_constr())
    _, posts = ml_oracle.read_trees([T, T2])
    want = np.asarray(ml_oracle.ultrametric_constr(posts, s.num_edges + 2))
    assert np.array_equal(M, want)
    # an ultrametric assignment satisfies every row
    x = np.zeros(s.num_edges + 2)
    idx = 0
    height = {}
    for t in s.trees:
        for v in t.traverse_postorder():
            height[id(v)] = 0.0 if v.is_leaf() else max(height[id(c)] for c in v.children) + 1.0
    for t in s.trees:
        for v in t.traverse_postorder():
            x[idx] = (height[id(v.parent)] if v.parent is not None else height[id(v)] + 1.0) - height[id(v)]
            idx += 1
    # equal root heights across trees: top up each root edge to a common level
    tops = []
    idx = 0
    for t in s.trees:
        n = sum(1 for _ in t.traverse_postorder())
        tops.append((idx + n - 1, height[id(t.root)]))
        idx += n
    H = max(h for _, h in tops) + 1.0
    for i, h in tops:
        x[i] = H - h
    assert np.allclose(M @ x, 0.0)


def test_ml_solver_host_surface():
    T = "((a:1,b:1,c:0.5):1,d:1):1;"
    msa = {"a": [1, -1], "b": [1, "?"], "c": [0, 2], "d": [-1, 0]}
    Q = [{1: 2.0, 2: 2.0}, {1: 1.0, 2: 3.0}]
    s = ML_solver([T], {"charMtrx": msa}, {"Q": Q})
    assert s.solver_name == "ML_solver"
    assert s.get_params() == {"phi": EPS, "nu": EPS}
    assert s.num_edges == 6 and s.has_polytomy
    assert s.Q[1] == {1: 0.25, 2: 0.75}
    assert s._leaf_row("a") == [1, "?"] and s._leaf_row("d") == ["?", 0]
    b = s.get_bound(fixed_nu=0.3)
    assert list(b.lb[:6]) == [s.dmin] * 6 and list(b.ub[:6]) == [s.dmax] * 6
    assert b.lb[6] == 0.3 - EPS and b.ub[6] == 0.3 + EPS
    assert b.lb[7] == EPS and b.ub[7] == 0.99
    s.x2params([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8], fixed_phi=0.05)
    assert [v.edge_length for v in s.trees[0].traverse_postorder()] == [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]
    assert s.get_params() == {"phi": 0.05, "nu": 0.7}
    with pytest.raises(ValueError):
        s.ultrametric_constr()   # the reference's constraint needs binary nodes


@pytest.mark.parametrize("seed", range(4))
def test_ml_oracle_mp_matches_double_where_nothing_cancels(seed):
    rng = random.Random(800 + seed)
    T, msa, Q = random_multifurcating(rng, 12, 5, 3, max_arity=4)
    nu, phi = 0.2 + 0.2 * seed, 0.1 + 0.05 * seed
    a = ml_oracle.negative_llh(T, msa, Q, nu, phi)
    b = ml_oracle.negative_llh_mp(T, msa, Q, nu, phi)
    assert math.isclose(a, b, rel_tol=1e-12)


def test_ml_oracle_double_loses_digits_at_tiny_nu():
    # the reference's 1 - p**nu at nu * d ~ 1e-12: the double-precision
    # restatement drifts from the exact value, which the device is held to
    T, msa, Q = random_multifurcating(random.Random(600), 10, 12, 3, max_arity=2, p_sil=0.0)
    a = ml_oracle.negative_llh(T, msa, Q, 1e-10, 1e-10)
    b = ml_oracle.negative_llh_mp(T, msa, Q, 1e-10, 1e-10)
    assert 1e-10 < abs(a - b) / abs(b) < 1e-6
