"""Topology search on the CPU: the oracle's optima over every rooted topology
against the reference's precomputed brute-force values, and the NNI search's
host logic (move listing, fixed-branch anchors, annealed acceptance, polytomy
marking) driven through the oracle-backed solver surface."""
import random
from copy import deepcopy

import pytest

from helpers import OracleSolver, load_topology_kats, rooted_topologies
from laml_amd.solvers import DEFAULT_STRATEGY
from laml_amd.topology_search import Topology_search, Topology_search_batched
from laml_amd.tree import read_tree_newick
from oracle.em_oracle import OracleEM

Q, MSA, BF, CASES = load_topology_kats()
PLACES = 4  # assertAlmostEqual(places=4) of the reference's tests


def _canon(nwk):
    t = read_tree_newick(nwk)

    def rec(v):
        if v.is_leaf():
            return v.label
        return tuple(sorted((rec(c) for c in v.children), key=str))
    return rec(t.root)


def test_rooted_topology_enumeration():
    for n, want in [(3, 3), (4, 15), (5, 105), (6, 945)]:
        topos = rooted_topologies([chr(ord('a') + i) for i in range(n)])
        assert len(topos) == want
        assert len({_canon(t) for t in topos}) == want


@pytest.mark.parametrize("key,ultra", [("msa1", False), ("msa2", False), ("msa2", True)])
def test_oracle_brute_force_matches_reference_optimum(key, ultra):
    best = min(OracleEM([t], MSA[key], Q).optimize(initials=1, ultra_constr=ultra)[0]
               for t in rooted_topologies(list("abcd")))
    want = BF[key + ("_ultra" if ultra else "")]
    assert round(best - want, PLACES) == 0, (best, want)


def test_oracle_score_of_fixed_topology():
    case = next(c for c in CASES if c["name"] == "test_5")
    nllh = OracleEM([case["T0"]], MSA["msa1"], Q).optimize(initials=1)[0]
    assert round(nllh - case["nllh"], PLACES) == 0


def test_oracle_fixed_branches_are_respected():
    # lengths of fixed branches stay (clamped at 2*dmin as in ML_solver.ini_brlens)
    t = "((a:0.3,c:0.2):0.1,(b:0.4,d:0.5):0.6):0.2;"
    o = OracleEM([t], MSA["msa1"], Q)
    fixed = [{('a', 'a'): 0.3, ('b', 'd'): 0.6}]
    o.optimize(initials=1, fixed_brlen=fixed, fixed_nu=0.1, fixed_phi=0.05)
    lens = {}
    from oracle.em_oracle import postorder
    for v in postorder(o.trees[0]):
        if v.label == 'a':
            lens['a'] = v.edge_length
        if v.a["fixed"]:
            lens.setdefault('fixed', []).append(v.edge_length)
    assert lens['a'] == 0.3 and sorted(lens['fixed']) == [0.3, 0.6]
    assert o.nu == 0.1 and o.phi == 0.05


def _searcher(cls, case, **kw):
    return cls([case["T0"]], OracleSolver, data={'charMtrx': MSA[case["msa"]]}, prior={'Q': Q},
               params={'nu': 0, 'phi': 0}, **kw)


def _strategy(case):
    st = deepcopy(DEFAULT_STRATEGY)
    st.update(case["strategy"])
    return st


@pytest.mark.parametrize("name", ["test_1", "test_3", "test_4", "test_6", "test_7", "test_8"])
def test_nni_search_host_logic_reaches_reference_optimum(name):
    case = next(c for c in CASES if c["name"] == name)
    random.seed(7)
    ts = _searcher(Topology_search, case)
    trees, score, params = ts.search(maxiter=200, nreps=1, strategy=_strategy(case))
    assert round(-score - case["nllh"], PLACES) == 0, (name, -score, case["nllh"], trees)
    assert ts.treeTopoList == trees and set(params) == {'nu', 'phi'}


def test_list_all_nni_moves():
    nwk = "(((a:1,b:1):1,(c:1,d:1):1):1,((e:1,f:1):1,g:1):1):1;"
    ts = Topology_search_batched([nwk], OracleSolver, data={'charMtrx': {}}, prior={'Q': []},
                                 params={'nu': 0.1, 'phi': 0.05})
    before = [t.newick() for t in ts.treeList_obj]
    random.seed(3)
    moves = ts.list_all_nni(DEFAULT_STRATEGY)
    n_int = sum(1 for t in ts.treeList_obj for v in t.traverse_postorder()
                if not v.is_leaf() and not v.is_root())
    assert len(moves) == 2 * n_int
    # the moves are undone (child order may differ: remove/append as in the reference)
    assert [_canon(t.newick()) for t in ts.treeList_obj] == [_canon(x) for x in before]
    canon0 = _canon(before[0])
    n_nodes = sum(1 for _ in ts.treeList_obj[0].traverse_postorder())
    seen = set()
    for nwks, st, (u, v, uc, w) in moves:
        c = _canon(nwks[0])
        assert c != canon0
        seen.add(c)
        assert st['fixed_nu'] == 0.1 and st['fixed_phi'] == 0.05
        assert len(st['fixed_brlen'][0]) == n_nodes - 5
        assert uc.parent is u and w.parent is v
        # every fixed anchor pair resolves to a node of the moved tree whose
        # length is the current one (all lengths are 1 here)
        t = read_tree_newick(nwks[0])
        from laml_amd.lca import find_LCAs
        for node in find_LCAs(t, list(st['fixed_brlen'][0].keys())):
            assert node.edge_length == 1
    assert len(seen) == len(moves)  # the moves are distinct topologies


def test_mark_polytomies_and_acceptance():
    ts = Topology_search(["((a,b),c,d,e);"], OracleSolver, data={'charMtrx': {}}, prior={'Q': []},
                         params={'nu': 0, 'phi': 0})
    ts.__mark_polytomies__()
    assert ts.has_polytomy
    marked = [v for t in ts.treeList_obj for v in t.traverse_preorder() if v.mark]
    assert len(marked) == 2 and all(v.edge_length == 1e-3 for v in marked)
    # annealing: T(0) = 1, T(T_cooldown) = 0 (only improvements accepted)
    assert abs(ts.a + ts.b - 1) < 1e-12
    assert abs(ts.a * ts.alpha_cooldown ** ts.T_cooldown + ts.b) < 1e-12
    random.seed(1)
    assert ts.__accept_proposal__(-10.0, -9.0, 0)
    assert not any(ts.__accept_proposal__(-10.0, -10.5, 25) for _ in range(100))
    random.seed(1)
    acc = sum(ts.__accept_proposal__(-10.0, -10.5, 0) for _ in range(4000)) / 4000
    assert abs(acc - 2.718281828459045 ** -0.5) < 0.03
