"""fastEM encodings (fastEM_solver.parse_tree / parse_data, phylogeny.py
build_phylogeny) on hand-worked examples: node relabelling, synthetic-root
removal, leaf ordering, state recoding, prior restriction.

parse_tree numbers nodes in treeswift's preorder, which reaches a node's
children from the last to the first (the order of the reference's own
annotation files, e.g. case_tests/test1/example3_annotations.txt: I0, I1,
c, I2, b, a for ((a,b),c)); for ((a,b),(c,d)) that is root, (c,d), d, c,
(a,b), b, a."""
import random

import numpy as np
import pytest

from laml_amd.fastem import parse_data, parse_tree
from laml_amd.tree import read_tree_newick


def test_parse_tree_relabels_leaves_first_then_internal_preorder():
    t = read_tree_newick("((a:1,b:2):3,(c:4,d:5):6);")
    random.seed(5)
    want_root = random.uniform(0.01, 0.5)
    random.seed(5)
    p = parse_tree(t)
    # leaves d, c, b, a; internal node_0 (root), node_1 = (c,d), node_4 = (a,b)
    assert p["relabeling_vector"] == ["d", "c", "b", "a", "node_0", "node_1", "node_4"]
    assert p["num_leaves"] == 4
    np.testing.assert_array_equal(p["parent"], [5, 5, 6, 6, -1, 4, 4])
    # children in the order the preorder reaches them
    assert p["children"][4] == [5, 6] and p["children"][5] == [0, 1] and p["children"][6] == [2, 3]
    assert p["branch_lengths"].dtype == np.float32
    np.testing.assert_array_equal(p["branch_lengths"], np.float32([5, 4, 2, 1, want_root, 6, 3]))
    assert p["branch_mask"].all()
    # internal labels are written back to the tree
    assert [v.label for v in t.traverse_preorder()] == ["node_0", "node_1", "d", "c", "node_4", "b", "a"]


def test_parse_tree_drops_synthetic_root_and_follows_leaf_order():
    t = read_tree_newick("(((a:1,b:2)x:3,(c:4,d:5)y:6)r:0.5);")
    for v in t.traverse_postorder():
        v.mark_fixed = v.label in ("a", "y")
    p = parse_tree(t, has_branch_mask=True, ordered_leaf_labels=["d", "c", "b", "a"])
    assert p["relabeling_vector"] == ["d", "c", "b", "a", "r", "y", "x"]
    np.testing.assert_array_equal(p["parent"], [5, 5, 6, 6, -1, 4, 4])
    np.testing.assert_array_equal(p["branch_lengths"], np.float32([5, 4, 2, 1, 0.5, 6, 3]))
    np.testing.assert_array_equal(p["branch_mask"], [True, True, True, False, True, False, True])
    if "nxtree" in p:
        g = p["nxtree"]
        assert sorted(g.successors(4)) == [5, 6] and g.nodes[0]["label"] == "d"


def test_parse_data_recodes_states_and_restricts_priors():
    t = read_tree_newick("((a:1,b:2):3,(c:4,d:5):6);")
    p = parse_tree(t)
    msa = {"a": [1, "?", 0], "b": [3, 2, 0], "c": [1, -1, 5], "d": [0, 2, 0]}
    Q = [{1: 0.2, 3: 0.5, 7: 0.3}, {2: 1.0}, {5: 0.6, 9: 0.4}]
    phy = parse_data(p, {"charMtrx": msa}, {"Q": Q})
    assert phy.num_leaves == 4 and phy.num_characters == 3 and phy.max_alphabet_size == 2 and phy.root == 4
    assert phy.character_matrix.dtype == np.int32
    # rows in leaf order d, c, b, a
    np.testing.assert_array_equal(phy.character_matrix, [[0, 1, 0], [1, -1, 1], [2, 1, 0], [1, -1, 0]])
    assert phy.mutation_priors.dtype == np.float32
    np.testing.assert_array_equal(phy.mutation_priors, np.float32([[0.2, 0.5], [1.0, 0.0], [0.6, 0.0]]))
    assert phy.state_mappings[0] == {1: 1, 2: 3}


def test_parse_data_missing_prior_is_an_error():
    t = read_tree_newick("(a:1,b:2);")
    p = parse_tree(t)
    with pytest.raises(KeyError):
        parse_data(p, {"charMtrx": {"a": [4], "b": [0]}}, {"Q": [{1: 1.0}]})
