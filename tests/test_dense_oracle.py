"""The dense per-state restatement (oracle/dense.py, the Felsenstein form of
the reference's fastEM recursions) against the compressed restatement of
EM_solver and against brute-force enumeration, including priors that do not
sum to one over a site's alphabet (the fastEM model's observed-state
alphabets)."""
import math
import random

import numpy as np
import pytest

from helpers import kat_inputs, load_kats
from oracle import brute
from oracle.dense import dense_estep
from oracle.em_oracle import OracleEM, postorder
from test_oracle import _random_tiny_case

KATS = load_kats()


@pytest.mark.parametrize("case", KATS["nllh_kats"][::3], ids=lambda c: c["name"])
def test_dense_equals_compressed_on_reference_constants(case):
    T, msa, Q, nu, phi = kat_inputs(case)
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    want = o.negative_llh()
    nll, _, _ = dense_estep(T, msa, o.Q, nu, phi)
    assert math.isclose(nll, want, rel_tol=1e-12, abs_tol=1e-12)
    assert abs(nll - case["nllh"]) < 2e-9


@pytest.mark.parametrize("seed", range(3))
def test_dense_equals_compressed_counts(seed):
    from laml_amd import sim
    tree, Q, codes, _ = sim.make_problem(40 + 20 * seed, 10, 3, 12, 1.0, 0.15, 0.2, 70 + seed)
    T = tree.newick()
    msa = sim.codes_to_charmtrx(tree, codes)
    o = OracleEM([T], msa, Q, nu=0.15, phi=0.2)
    o.az_partition()
    o.estep()
    want = np.array([[sum(v.a["S"][j]) for j in range(5)] for v in postorder(o.trees[0])])
    _, Rt = o.phi_terms()
    nll, counts, rt = dense_estep(T, msa, o.Q, 0.15, 0.2)
    assert math.isclose(nll, o.negative_llh(), rel_tol=1e-12)
    np.testing.assert_allclose(counts, want, rtol=1e-9, atol=1e-11)
    assert math.isclose(rt, Rt, rel_tol=1e-10)


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("minus_one_missing", [False, True])
def test_dense_equals_brute_force_with_unnormalised_priors(seed, minus_one_missing):
    rng = random.Random(100 + seed)
    n = 3 + seed % 3
    T, msa, Q = _random_tiny_case(rng, n, 3, 2 if n == 5 else 3)
    # observed-state alphabets with their raw prior mass (sums below one)
    Q = [{a: p * rng.uniform(0.3, 0.9) for a, p in q.items()} for q in Q]
    nu = [0.05, 0.3, 0.9, 0.0][seed % 4]
    phi = [0.1, 0.35, 0.0][seed % 3]
    if minus_one_missing:
        dense_kw = dict(missing=("?", -1), silenced=())
        msa_b = {c: ["?" if x == -1 else x for x in row] for c, row in msa.items()}
    else:
        dense_kw = {}
        msa_b = msa
    nll, counts, rt = dense_estep(T, msa, Q, nu, phi, **dense_kw)
    tot, bt_rt = 0.0, 0.0
    want = None
    from oracle.em_oracle import parse_newick
    labels = [v.label for v in postorder(parse_newick(T))]
    want = np.zeros((len(labels), 5))
    for k in range(len(Q)):
        col = {c: row[k] for c, row in msa_b.items()}
        L, cnt, r = brute.enumerate_site(T, col, Q[k], nu, phi)
        if L == 0.0:
            pytest.skip("data impossible under these parameters")
        tot += math.log(L)
        bt_rt += r
        for i, lab in enumerate(labels):
            want[i] += cnt[lab]
    assert math.isclose(nll, -tot, rel_tol=1e-12, abs_tol=1e-12)
    np.testing.assert_allclose(counts, want, rtol=1e-9, atol=1e-12)
    assert math.isclose(rt, bt_rt, rel_tol=1e-9, abs_tol=1e-12)
