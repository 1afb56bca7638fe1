"""fastEM encodings (fastEM_solver.parse_tree / parse_data, phylogeny.py
build_phylogeny) on hand-worked examples: node relabelling, synthetic-root
removal, leaf ordering, state recoding, prior restriction."""
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
    assert p["relabeling_vector"] == ["a", "b", "c", "d", "node_0", "node_1", "node_4"]
    assert p["num_leaves"] == 4
    np.testing.assert_array_equal(p["parent"], [5, 5, 6, 6, -1, 4, 4])
    assert p["children"][4] == [5, 6] and p["children"][5] == [0, 1]
    assert p["branch_lengths"].dtype == np.float32
    np.testing.assert_array_equal(p["branch_lengths"], np.float32([1, 2, 4, 5, want_root, 3, 6]))
    assert p["branch_mask"].all()
    # internal labels are written back to the tree
    assert [v.label for v in t.traverse_preorder()] == ["node_0", "node_1", "a", "b", "node_4", "c", "d"]


def test_parse_tree_drops_synthetic_root_and_follows_leaf_order():
    t = read_tree_newick("(((a:1,b:2)x:3,(c:4,d:5)y:6)r:0.5);")
    for v in t.traverse_postorder():
        v.mark_fixed = v.label in ("a", "y")
    p = parse_tree(t, has_branch_mask=True, ordered_leaf_labels=["d", "c", "b", "a"])
    assert p["relabeling_vector"] == ["d", "c", "b", "a", "r", "x", "y"]
    np.testing.assert_array_equal(p["parent"], [6, 6, 5, 5, -1, 4, 4])
    np.testing.assert_array_equal(p["branch_lengths"], np.float32([5, 4, 2, 1, 0.5, 3, 6]))
    np.testing.assert_array_equal(p["branch_mask"], [True, True, True, False, True, True, False])
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
    np.testing.assert_array_equal(phy.character_matrix, [[1, -1, 0], [2, 1, 0], [1, -1, 1], [0, 1, 0]])
    assert phy.mutation_priors.dtype == np.float32
    np.testing.assert_array_equal(phy.mutation_priors, np.float32([[0.2, 0.5], [1.0, 0.0], [0.6, 0.0]]))
    assert phy.state_mappings[0] == {1: 1, 2: 3}


def test_parse_data_missing_prior_is_an_error():
    t = read_tree_newick("(a:1,b:2);")
    p = parse_tree(t)
    with pytest.raises(KeyError):
        parse_data(p, {"charMtrx": {"a": [4], "b": [0]}}, {"Q": [{1: 1.0}]})
