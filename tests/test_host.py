"""CPU tests of the host side: tree handling, forest layout (integer work,
bit-exact), the C ABI's exported symbols and its host-side validation."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import load_test4
from laml_amd import _lib, sim
from laml_amd.forest import ForestLayout, SiteTable, tree_arrays_from_tree
from laml_amd.lca import find_LCAs
from laml_amd.tree import read_tree_newick
from oracle.em_oracle import OracleEM, parse_newick, postorder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    header = open(os.path.join(ROOT, "include", "laml_pmm.h")).read()
    names = set(re.findall(r"^(?:int|int32_t|int64_t|void|const char\*)\s+(laml_\w+)\(", header, re.M))
    assert names, "no declarations parsed"
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)
    for n in names:
        assert hasattr(lib, n)
    assert lib.laml_version().startswith(b"laml-amd")


def test_newick_roundtrip_and_order():
    s = "((a:1,b:0.5)x:0.25,(c:2,d:1e-05)y:1)r:0.1;"
    t = read_tree_newick(s)
    assert [v.label for v in t.traverse_postorder()] == ["a", "b", "x", "c", "d", "y", "r"]
    assert t.newick() == "((a:1,b:0.5)x:0.25,(c:2,d:1e-05)y:1)r:0.1;"
    # treeswift's preorder (and so its leaf order) reaches the last child first
    assert [v.label for v in t.traverse_preorder()] == ["r", "y", "d", "c", "x", "b", "a"]
    assert [v.label for v in t.traverse_leaves()] == ["d", "c", "b", "a"]


def test_tree_matches_oracle_parser():
    T, msa, Q = load_test4()
    t = read_tree_newick(T)
    t.suppress_unifurcations()
    from oracle.em_oracle import suppress_unifurcations
    o = suppress_unifurcations(parse_newick(T))
    a = [(v.label, v.edge_length) for v in t.traverse_postorder()]
    b = [(v.label, v.edge_length) for v in postorder(o)]
    assert a == b


def test_suppress_unifurcations():
    t = read_tree_newick("(((a:1,b:1)x:1)u:2,c:1)r;")
    t.suppress_unifurcations()
    assert t.newick() == "(c:1,(a:1,b:1)x:3)r;"
    t = read_tree_newick("((a:1,b:1)x:1)r:2;")
    t.suppress_unifurcations()
    assert t.newick() == "(a:1,b:1)x:3;"


def test_resolve_polytomies_binary_and_zero_length():
    t = read_tree_newick("(a:1,b:1,c:1,d:1)r;")
    new = t.resolve_polytomies()
    assert len(new) == 2
    assert all(len(v.children) in (0, 2) for v in t.traverse_postorder())
    assert all(v.edge_length == 0.0 for v in new)
    assert sorted(v.label for v in t.traverse_leaves()) == ["a", "b", "c", "d"]


def test_collapse_short_branches():
    # x is contracted: its children join r's children at the end and keep
    # their lengths (what the reference's collapsed-tree output shows)
    t = read_tree_newick("((a:1,b:1)x:0.00001,c:1)r;")
    t.collapse_short_branches(0.00005)
    assert t.newick() == "(c:1,a:1,b:1)r:0;"


def test_find_lcas():
    t = read_tree_newick("((a,b)x,((c,d)y,e)z)r;")
    res = find_LCAs(t, [("a", "b"), ("c", "e"), ("a", "e"), ("d", "d"), ("q", "c")])
    assert [v.label for v in res] == ["x", "z", "r", "d", "c"]


def _layout_for(T, msa, Q):
    t = read_tree_newick(T)
    t.suppress_unifurcations()
    sites = SiteTable(Q)
    ta = tree_arrays_from_tree(t, lambda v: sites.encode_row(msa[v.label]))
    from laml_amd.forest import ProblemSpec

    return ForestLayout([ProblemSpec([ta], sites, 1, 0.01)]), t, sites


def test_layout_levels_and_validation():
    T, msa, Q = load_test4()
    lay, t, sites = _layout_for(T, msa, Q)
    lay.validate()
    N = lay.num_nodes
    # children strictly below, leaves first, parent pointers consistent
    lvl = np.zeros(N, dtype=int)
    for h in range(lay.num_levels):
        lvl[lay.level_off[h]:lay.level_off[h + 1]] = h
    for v in range(lay.num_leaves, N):
        for c in (lay.child0[v], lay.child1[v]):
            assert lvl[c] < lvl[v] and lay.parent[c] == v
            assert lvl[v] == 1 + max(lvl[lay.child0[v]], lvl[lay.child1[v]])
    assert (lay.child0[: lay.num_leaves] == -1).all()
    # leaf codes in level order decode back to the character matrix
    K = sites.K
    inv_maps = [{j: s for s, j in m.items()} for m in sites.maps]
    nodes = list(t.traverse_postorder())
    for i, g in enumerate(lay.order[: lay.num_leaves]):
        v = nodes[g]
        row = lay.codes[i * K:(i + 1) * K]
        dec = [0 if c == 0 else ("?" if c == -1 else (-1 if c == -2 else inv_maps[k][int(c)])) for k, c in
               enumerate(row)]
        assert dec == msa[v.label]


def test_validation_rejects_bad_description():
    T, msa, Q = load_test4()
    lay, _, _ = _layout_for(T, msa, Q)
    lay.child0[-1] = lay.child1[-1]
    with pytest.raises(_lib.LamlError):
        lay.validate()
    lay, _, _ = _layout_for(T, msa, Q)
    lay.codes = lay.codes.copy()
    lay.codes[3] = 999
    with pytest.raises(_lib.LamlError, match="leaf code"):
        lay.validate()


def test_prior_with_mass_on_unedited_state_rejected():
    with pytest.raises(ValueError):
        SiteTable([{0: 0.2, 1: 0.8}])


def test_unknown_observed_state_rejected():
    sites = SiteTable([{1: 1.0}])
    with pytest.raises(KeyError):
        sites.encode_row([7])


def test_synthetic_generator_semantics():
    # proportions of unedited / missing leaf entries follow the model
    tree, Q, codes, truth = sim.make_problem(2000, 60, 5, 20, 1.0, 0.1, 0.2, 5)
    assert codes.shape == (2000, 60)
    depth_len = 1.0  # stem + root-to-tip
    p0 = np.exp(-(1 + 0.1) * depth_len)  # P(still unedited at a leaf, no dropout)
    frac0 = (codes == 0).mean()
    assert abs(frac0 - p0 * 0.8) < 0.02
    p_sil = 1 - np.exp(-0.1 * depth_len)
    frac_miss = (codes == -1).mean()
    assert abs(frac_miss - (p_sil + (1 - p_sil) * 0.2)) < 0.02
    # same seed, same data
    _, _, codes2, _ = sim.make_problem(2000, 60, 5, 20, 1.0, 0.1, 0.2, 5)
    assert (codes == codes2).all()


def test_yule_tree_is_ultrametric():
    rng = np.random.Generator(np.random.PCG64(3))
    t = sim.yule_tree(500, rng)
    tr = t.to_tree()
    dist = {}
    for v in tr.traverse_preorder():
        dist[v] = (v.edge_length) + (dist[v.parent] if v.parent is not None else 0.0)
    leaves = [dist[v] for v in tr.traverse_leaves()]
    assert max(leaves) - min(leaves) < 1e-12
    assert abs(max(leaves) - 1.0) < 1e-12


def test_ultrametric_start_heights_are_strictly_feasible():
    """Random initial lengths (ML_solver.ini_brlens draws up to dmax/2 per
    branch) can make the plain bottom-up placement exceed dmax on a branch to
    a pinned leaf; the start construction must still be strictly feasible."""
    import random
    from helpers import load_topology_kats, rooted_topologies
    from laml_amd.solvers import EM_solver
    from laml_amd.ultra import HeightModel
    Q, MSA, _, _ = load_topology_kats()
    bad_plain = 0
    for i, T in enumerate(rooted_topologies(list("abcde"))):
        s = EM_solver([T], {'charMtrx': {**MSA["msa2"], 'e': [1, 0, 1, 0, 1]}}, {'Q': Q}, {'nu': 0, 'phi': 0})
        random.seed(i)
        s.x2params(s.ini_all())
        L = ForestLayout([s._problem_spec(3)])
        d = L.lengths_to_level([np.asarray([v.edge_length for v in t.traverse_postorder()]) for t in s.trees])
        M = HeightModel(L, d)
        t0 = M.heights_from_lengths(d, L.dmin, L.dmax)
        bad_plain += not M.check_strict(t0, L.free, L.dmin, L.dmax)
        t = M.start_heights(d, L.free, L.dmin, L.dmax)
        assert M.check_strict(t, L.free, L.dmin, L.dmax)
        lens = M.lengths(t, L.num_nodes)
        # ultrametric: every leaf-to-root path has the same length
        par = L.parent
        depth = []
        for leaf in range(L.num_leaves):
            x, acc = leaf, 0.0
            while x >= 0:
                acc += lens[x]
                x = par[x]
            depth.append(acc)
        assert max(depth) - min(depth) < 1e-9
    assert bad_plain > 0  # the case the scaled placement exists for does occur


def test_chain_byte_accounting():
    """The benchmark's algorithmic bytes with outside chains: every heavy
    child below the split saves its outside-row write and read (76 B per
    site); without a chain region the count is the level pass's."""
    import bench

    probs = [sim.make_problem(300 + 50 * i, 40, 3, 10, 1.0, 0.1, 0.1, 60 + i) for i in range(3)]
    F = ForestLayout([sim.problem_spec(t, q, c) for (t, q, c, _) in probs])
    base = bench.algorithmic_bytes(F)
    assert bench.algorithmic_bytes(F, chain_split=1) == base
    n, N = F.num_leaves, F.num_nodes
    # leaf counts by recursion from each root
    nl = {}

    def count(v):
        if v < n:
            return 1
        if v not in nl:
            nl[v] = count(int(F.child0[v])) + count(int(F.child1[v]))
        return nl[v]

    heavy = bench.heavy_children(F)
    K = F.tree_K[F.tree_of]
    for v in range(n, N):
        a, b = int(F.child0[v]), int(F.child1[v])
        want = [c for c in (a, b) if c >= n]
        if len(want) == 2:
            want = [b] if count(b) > count(a) else [a]
        assert heavy[v] == (want[0] if want else -1)
    H = F.num_levels
    for split in (H, 5, 2):
        B = int(F.level_off[split])
        saved = 76 * sum(int(K[heavy[v]]) for v in range(n, B) if heavy[v] >= 0)
        ins, out = bench.algorithmic_bytes(F, chain_split=split)
        assert ins == base[0] and out == base[1] - saved
        ins, out = bench.algorithmic_bytes(F, in_chain_split=split)
        assert out == base[1] and ins == base[0] - saved // 76 * 22


def test_pipelined_batch_split():
    """Contiguous, non-empty, balanced groups of problems for the streams."""
    from laml_amd.streams import split_balanced

    probs = [sim.make_problem(n, K, 3, 10, 1.0, 0.1, 0.1, 70 + i)
             for i, (n, K) in enumerate([(300, 40), (50, 40), (50, 40), (600, 40), (40, 10), (200, 80)])]
    specs = [sim.problem_spec(t, q, c) for (t, q, c, _) in probs]
    for parts in (1, 2, 3, 6, 9):
        g = split_balanced(specs, parts)
        assert len(g) == min(parts, len(specs))
        assert [i for grp in g for i in grp] == list(range(len(specs)))
        assert all(grp for grp in g)
    w = [sum(t.n_nodes for t in p.trees) * p.sites.K for p in specs]
    g = split_balanced(specs, 2)
    halves = [sum(w[i] for i in grp) for grp in g]
    assert max(halves) <= sum(w) - min(w)
