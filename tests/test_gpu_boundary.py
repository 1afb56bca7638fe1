"""The Python boundary as the reference's parallel search uses it.

- A solver that has scored (ultrametric model attached, device forest
  built) pickles without any device state and, unpickled, scores the same
  again, bit for bit (Topology_search_parallel.py:25-26,62-63 ship solvers
  between processes; the unpickled copy rebuilds its forest on first use).
- Batched candidate scoring leaves the caller's random generator as it was,
  as the reference's forked pool workers do, so the search's own draws do
  not depend on which scoring path ran (the quartet path draws nothing).
"""
import pickle
import random

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.solvers import DEFAULT_STRATEGY, EM_solver
from laml_amd.topology_search import Topology_search_batched, score_batch

pytestmark = pytest.mark.gpu


def _problem(n=60, K=25, seed=4):
    tree, Q, codes, _ = sim.make_problem(n, K, 3, 12, 1.0, 0.1, 0.1, seed)
    return tree.newick(), sim.codes_to_charmtrx(tree, codes), Q


def test_scored_solver_pickles_and_rescores_equal():
    T, msa, Q = _problem()
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"nu": 0.1, "phi": 0.1})
    st = dict(DEFAULT_STRATEGY, ultra_constr=True, fixed_brlen=None)
    random.seed(11)
    s.score_tree(strategy=st)
    assert s._forest is None or s._forest.handle is not None
    blob = pickle.dumps(s)
    s2 = pickle.loads(blob)
    assert s2._forest is None and not hasattr(s2, "_last_estep")
    for v in s2.__dict__.values():
        assert not isinstance(v, torch.Tensor)
    random.seed(12)
    a = s.score_tree(strategy=st)
    random.seed(12)
    b = s2.score_tree(strategy=st)
    print("rescored", a, b, flush=True)
    assert a == b
    assert s.get_tree_newick() == s2.get_tree_newick() and s.get_params() == s2.get_params()


def test_batched_scoring_leaves_the_random_generator_unchanged():
    T, msa, Q = _problem(40, 20, 8)
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"nu": 0.1, "phi": 0.1})
    random.seed(3)
    s.score_tree(strategy=dict(DEFAULT_STRATEGY, fixed_brlen=None))
    ts = Topology_search_batched(s.get_tree_newick(), EM_solver, data={"charMtrx": msa}, prior={"Q": Q},
                                 params=s.get_params())
    random.seed(5)
    moves = ts.list_all_nni(dict(DEFAULT_STRATEGY, local_brlen_opt=False))[:6]
    state = random.getstate()
    res = score_batch([(m[0], m[1]) for m in moves], ts.data, ts.prior, ts.params, solver_cls=EM_solver)
    assert random.getstate() == state
    assert all(r[2] == "optimal" for r in res)
    # every candidate drew its seed from the caller's state, as a candidate
    # alone in a freshly forked worker does
    want = int(random.random() * 10000)
    assert all(r[0].em_seed == want for r in res)
