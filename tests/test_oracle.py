"""Pins the CPU oracle before anything is checked against it.

- the reference's 21 NLL known-answer constants (unit_tests_EMSolver.py test_1..20, 46)
- brute-force enumeration of internal states (likelihood, edge posteriors S0..S4, dropout responsibilities)
- the reference's reduced-tree invariant for the outside pass (unit_tests_EMSolver.py:267-303)
- the reference's EM end-to-end KAT (unit_tests_EMSolver.py:535-568, test_47)
- optimality (KKT) of the oracle's exact M-step solutions
"""
import math
import random

import pytest

from helpers import kat_inputs, load_kats, load_test1, load_test4
from oracle import brute
from oracle.em_oracle import (OracleEM, branch_grad, branch_obj, nu_grad, nu_obj, parse_newick, postorder,
                              preorder, solve_branch, solve_nu)

KATS = load_kats()


@pytest.mark.parametrize("case", KATS["nllh_kats"], ids=lambda c: c["name"])
def test_reference_nllh_constants(case):
    T, msa, Q, nu, phi = kat_inputs(case)
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    got = o.negative_llh()
    # the reference checks 5 places; the constants were produced with
    # eps-sized (1e-10) parameters, so they agree here to ~1e-9
    assert abs(got - case["nllh"]) < 2e-9


@pytest.mark.parametrize("case", KATS["nllh_kats"], ids=lambda c: c["name"])
def test_brute_force_matches_reference_constants(case):
    T, msa, Q, nu, phi = kat_inputs(case)
    assert abs(brute.brute_nllh(T, msa, Q, nu, phi) - case["nllh"]) < 2e-9


def _random_tiny_case(rng, n_leaves, K, m):
    # random binary tree with labelled internal nodes
    nodes = ["L%d" % i for i in range(n_leaves)]
    subs = ["%s:%.3f" % (x, rng.uniform(0.05, 1.5)) for x in nodes]
    ii = 0
    while len(subs) > 1:
        i, j = rng.sample(range(len(subs)), 2)
        a, b = subs[i], subs[j]
        for idx in sorted((i, j), reverse=True):
            subs.pop(idx)
        subs.append("(%s,%s)I%d:%.3f" % (a, b, ii, rng.uniform(0.05, 1.5)))
        ii += 1
    T = subs[0] + ";"
    Q = []
    for _ in range(K):
        w = [rng.random() + 0.05 for _ in range(m)]
        s = sum(w)
        Q.append({a + 1: w[a] / s for a in range(m)})
    choices = [0, "?", -1] + list(range(1, m + 1))
    msa = {x: [rng.choice(choices) for _ in range(K)] for x in nodes}
    return T, msa, Q


@pytest.mark.parametrize("seed", range(12))
def test_oracle_matches_brute_force_posteriors(seed):
    rng = random.Random(seed)
    n = 3 + seed % 3
    T, msa, Q = _random_tiny_case(rng, n, 3, 2 if n == 5 else 3)
    nu = [0.0, 0.05, 0.3, 0.9][seed % 4]
    phi = [0.0, 0.1, 0.35][seed % 3]
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    o.az_partition()
    o.estep()
    root = o.trees[0]
    R, Rt = o.phi_terms()
    bt_rt = 0.0
    for k in range(len(Q)):
        col = {c: row[k] for c, row in msa.items()}
        L, counts, rt = brute.enumerate_site(T, col, Q[k], nu, phi)
        if L == 0.0:  # data impossible under these parameters
            assert root.a["L0"][k] == -math.inf or math.isnan(root.a["L0"][k])
            continue
        bt_rt += rt
        assert math.isclose(math.log(L), root.a["L0"][k], rel_tol=1e-12, abs_tol=1e-12)
        for v in postorder(root):
            for j in range(5):
                assert abs(v.a["S"][j][k] - counts[v.label][j]) < 1e-9, (v.label, j, k)
    if all(math.isfinite(x) for x in root.a["L0"]):
        assert abs(Rt - bt_rt) < 1e-9


def _reduced_trees(root):
    """unit_tests_EMSolver.py:247-264: for every node x, the Newick of the tree
    with x's subtree replaced by a leaf named x."""
    from oracle.em_oracle import ONode

    def newick(v):
        if not v.children:
            s = v.label
        else:
            s = "(" + ",".join(newick(c) for c in v.children) + ")" + (v.label or "")
        if v.edge_length is not None:
            s += ":" + repr(v.edge_length)
        return s

    out = {}
    for x in preorder(root):
        saved = x.children
        x.children = []
        out[x.label] = newick(root) + ";"
        x.children = saved
    return out


def _check_outside_invariant(T, msa, Q, phi, nu, label_internal=False, places=9):
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    if label_internal:
        i = 0
        for v in preorder(o.trees[0]):
            if v.children:
                v.label = "I%d" % i
                i += 1
    o.az_partition()
    o.estep_in()
    o.estep_out()
    K = o.numsites
    red = _reduced_trees(o.trees[0])
    for v in postorder(o.trees[0]):
        for code, key, shift in ((0, "out0", math.log(1 - phi) if phi < 1 else 0.0), (-1, "out1", 0.0)):
            m = dict(msa)
            m[v.label] = [code] * K
            o2 = OracleEM([red[v.label]], m, Q, nu=nu, phi=phi)
            o2.az_partition()
            o2.estep_in()
            for k in range(K):
                want = o2.trees[0].a["L0"][k]
                got = v.a[key][k] + shift
                if want == -math.inf:
                    assert got == -math.inf
                else:
                    assert abs(want - got) < 10 ** (-places), (v.label, key, k, want, got)


@pytest.mark.parametrize("case", KATS["outside_invariants"], ids=lambda c: c["name"])
def test_outside_reduced_tree_invariant(case):
    T, msa, Q, nu, phi = kat_inputs(case)
    _check_outside_invariant(T, msa, Q, phi, nu)


def test_outside_reduced_tree_invariant_test1_file():
    T, msa, Q = load_test1()
    case = KATS["outside_invariant_file"]
    # a subset of the 299 nodes keeps this CPU test to seconds; the GPU tests
    # compare every node's counts against this oracle
    msa_small = msa
    o = OracleEM([T], msa_small, Q, nu=case["nu"], phi=case["phi"])
    assert o.numsites == 60
    _check_outside_invariant_subset(T, msa, Q, case["phi"], case["nu"], every=23)


def _check_outside_invariant_subset(T, msa, Q, phi, nu, every):
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    i = 0
    for v in preorder(o.trees[0]):
        if v.children:
            v.label = "I%d" % i
            i += 1
    o.az_partition()
    o.estep_in()
    o.estep_out()
    K = o.numsites
    red = _reduced_trees(o.trees[0])
    for idx, v in enumerate(postorder(o.trees[0])):
        if idx % every:
            continue
        for code, key, shift in ((0, "out0", math.log(1 - phi)), (-1, "out1", 0.0)):
            m = dict(msa)
            m[v.label] = [code] * K
            o2 = OracleEM([red[v.label]], m, Q, nu=nu, phi=phi)
            o2.az_partition()
            o2.estep_in()
            for k in range(K):
                want = o2.trees[0].a["L0"][k]
                got = v.a[key][k] + shift
                assert abs(want - got) < 1e-8 * max(1.0, abs(want))


def test_reference_em_kat_test47():
    em = KATS["em_kat"]
    T, msa, Q = load_test4()
    o = OracleEM([T], msa, Q, 0, 0)
    nllh, _, nu, phi, _, status = o.optimize(initials=1, random_seeds=em["random_seed"])
    assert abs(nllh - em["nllh"]) / em["nllh"] < em["nllh_rel_tol"]
    assert round(abs(phi - em["phi"]), em["param_places"]) == 0
    assert round(abs(nu - em["nu"]), em["param_places"]) == 0
    assert status == "optimal"


@pytest.mark.parametrize("seed", range(6))
def test_mstep_branch_solution_is_optimal(seed):
    rng = random.Random(seed)
    for _ in range(50):
        s = [rng.random() * rng.choice([1e-6, 1, 30]) for _ in range(5)]
        nu = rng.choice([0.0, 1e-6, 0.05, 0.5, 1.0])
        use = nu > 1e-5
        d = solve_branch(s, nu, use, use, 0.005, 10.0)
        g = branch_grad(s, nu, use, use, d)
        if 0.005 < d < 10.0:
            assert abs(g) < 1e-8 * (1 + abs(s[1] / d))
        elif d == 0.005:
            assert g <= 0
        else:
            assert g >= 0
        f = branch_obj(s, nu, use, use, d)
        for h in (1e-4, 1e-2):
            for dd in (d - h, d + h):
                if 0.005 <= dd <= 10.0:
                    assert branch_obj(s, nu, use, use, dd) <= f + 1e-12 * abs(f)


@pytest.mark.parametrize("seed", range(4))
def test_mstep_nu_solution_is_optimal(seed):
    rng = random.Random(100 + seed)
    rows = [[rng.random() * 10, rng.random() * 5, rng.random() * 0.5 + 1e-6, rng.random() * 3,
             rng.random() * 0.3 + 1e-6] for _ in range(40)]
    d = [rng.uniform(0.005, 2) for _ in rows]
    nu = solve_nu(rows, d)
    f = nu_obj(rows, d, nu)
    if 0 < nu < 1:
        assert abs(nu_grad(rows, d, nu)) < 1e-7 * sum(r[0] for r in rows)
    for dd in (nu * 0.99, nu * 1.01, min(1.0, nu + 1e-3)):
        if 0 < dd <= 1:
            assert nu_obj(rows, d, dd) <= f + 1e-12 * abs(f)


def test_branch_closed_form_when_no_silencing():
    # with nu = 0 the branch problem has the closed form of EM_solver.py:387-390
    s = [3.0, 1.5, 1e-6, 2.0, 1e-6]
    d = solve_branch(s, 0.0, False, False, 1e-9, 50.0)
    assert math.isclose(d, -math.log(s[0] / (s[0] + s[1])), rel_tol=1e-13)


def test_newick_parse_postorder():
    r = parse_newick("((a:1,b:2)x:3,(c:4,d:5)y:6)r:7;")
    assert [v.label for v in postorder(r)] == ["a", "b", "x", "c", "d", "y", "r"]
    assert [v.edge_length for v in postorder(r)] == [1, 2, 3, 4, 5, 6, 7]
