"""Quartet-local EM of NNI candidates (laml_quartet_em) against the CPU
oracle's EM of the same candidate, set up as the reference's apply_nni sets
it up with local_brlen_opt (Topology_search.py:247-306: the branches outside
the quartet fixed at the lengths their anchor pairs name, nu and phi fixed,
ini_brlens raising every branch to 2 dmin, EM_solver.EM_optimization on the
five free branches): the same NLL (1e-9 relative), the same five optimised
branch lengths (1e-6) and the same EM iteration count.  The device's full
per-candidate solver is held to the same three quantities."""
import math
import random

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.quartet import QuartetScorer, _clade_index
from laml_amd.solvers import DEFAULT_STRATEGY, EM_solver
from laml_amd.topology_search import Topology_search_batched, _nni_swap, _nni_unswap
from laml_amd.tree import read_tree_newick
from oracle.em_oracle import OracleEM, postorder
from oracle.pmm_ref import RefProblem

pytestmark = pytest.mark.gpu


def _search_state(n, K, seed):
    tree, Q, codes, _ = sim.make_problem(n, K, 3, 15, 1.0, 0.1, 0.15, seed)
    msa = sim.codes_to_charmtrx(tree, codes)
    random.seed(seed)
    s = EM_solver([tree.newick()], {"charMtrx": msa}, {"Q": Q}, {"nu": 0, "phi": 0})
    s.score_tree(strategy=dict(DEFAULT_STRATEGY, fixed_brlen=None))
    ts = Topology_search_batched(s.get_tree_newick(), EM_solver, data={"charMtrx": msa}, prior={"Q": Q},
                                 params=s.get_params())
    return ts, msa, Q


def _lengths_by_clade(nwk):
    t = read_tree_newick(nwk)
    idx, below = _clade_index(t)
    return {below[id(v)]: v.edge_length for v in t.traverse_postorder()}


@pytest.mark.parametrize("n,K,seed", [(40, 20, 3), (90, 70, 8)])
def test_quartet_em_equals_full_candidate_em(n, K, seed):
    ts, msa, Q = _search_state(n, K, seed)
    random.seed(1)
    moves = ts._list_moves()
    sel = moves[:: max(1, len(moves) // 14)]
    scorer = QuartetScorer(ts, DEFAULT_STRATEGY, EM_solver)
    res = scorer.score(sel)
    iters_q = scorer.iters.copy()
    below = _clade_index(ts.treeList_obj[0])[1]
    for i, (mv, (score, lazy)) in enumerate(zip(sel, res)):
        u, v, uc, w = mv
        st = ts._local_strategy(u, DEFAULT_STRATEGY)
        _nni_swap(u, v, uc, w)
        nwks = [t.newick() for t in ts.treeList_obj]
        _nni_unswap(u, v, uc, w)
        s = EM_solver(nwks, {"charMtrx": msa}, {"Q": Q}, ts.params)
        nllh, status = s.optimize(initials=1, verbose=-1, fixed_phi=st["fixed_phi"], fixed_nu=st["fixed_nu"],
                                  fixed_brlen=st["fixed_brlen"], random_seeds=[1234])
        # the oracle: the same candidate, marked and started as the reference does
        o = OracleEM(nwks, msa, Q)
        o.mark_fixed(st["fixed_brlen"])
        x0 = o.ini_all(random.Random(1234), fixed_phi=st["fixed_phi"], fixed_nu=st["fixed_nu"])
        o.x2params(x0, fixed_phi=st["fixed_phi"], fixed_nu=st["fixed_nu"])
        rp, d0 = RefProblem.from_oracle(o)
        onll, oit, od, _, _ = rp.em(d0, o.nu, o.phi, opt_nu=False, opt_phi=False)
        print("move", i, -score, nllh, onll, iters_q[i], s.em_iterations, oit, flush=True)
        assert math.isclose(-score, onll, rel_tol=1e-9), (i, -score, onll)
        assert math.isclose(nllh, onll, rel_tol=1e-9), (i, nllh, onll)
        assert int(iters_q[i]) == oit, (i, int(iters_q[i]), oit)
        assert s.em_iterations == oit, (i, s.em_iterations, oit)
        # the five free branches, by clade in the candidate tree
        A = next(x for x in u.child_nodes() if x is not uc)
        cA, cW, cC = below[id(A)], below[id(w)], below[id(uc)]
        clades = [cA, cW, cA | cW, cC, cA | cW | cC]
        got = lazy.lengths
        want = _lengths_by_clade(s.get_tree_newick()[0])
        oblow = {}
        oracle_len = {}
        for v, x in zip([v for r in o.trees for v in postorder(r)], od):
            cl = frozenset([v.label]) if not v.children else frozenset().union(*[oblow[id(c)] for c in v.children])
            oblow[id(v)] = cl
            oracle_len[cl] = x
        for j, cl in enumerate(clades):
            assert abs(got[j] - oracle_len[cl]) <= 1e-6 * max(1.0, oracle_len[cl]), (i, j, got[j], oracle_len[cl])
            assert abs(got[j] - want[cl]) <= 1e-6 * max(1.0, want[cl]), (i, j, got[j], want[cl])
        # the accepted-candidate tree text carries the same lengths
        mat = _lengths_by_clade(lazy.materialise().get_tree_newick()[0])
        for cl, x in want.items():
            if cl in mat:
                assert abs(mat[cl] - x) <= 1e-6 * max(1.0, x)
    scorer.close()
