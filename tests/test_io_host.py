"""Input/output formats (sequence_lib.py, runner.py:323-435) on the CPU:
readers against the reference's fixture files and hand-worked examples,
duplicate handling, and the tree/parameter writers."""
import json
import pickle

import pytest

from helpers import golden_path
from laml_amd import io
from laml_amd.tree import read_tree_newick
from oracle.em_oracle import read_charmtrx, read_priors_csv


def test_read_charmtrx_matches_the_fixture_reader():
    path = golden_path("test4_charMtrx.txt")
    D, names = io.read_sequences(path, filetype="charMtrx", delimiter=",", masked_symbol="?")
    assert D == read_charmtrx(path, masked_symbol="?")
    assert len(names) == len(next(iter(D.values())))


def test_missing_symbol_rules(tmp_path):
    p = tmp_path / "m.csv"
    p.write_text("cell_name,a,b,c,d,e,f,g,h\nx,-,?,-1,N,-3,5,0,abc\ny,1,2,3,N,4,5,6,7\n")
    D, names = io.read_sequences(str(p), masked_symbol="N")
    assert names == list("abcdefgh")
    assert D["x"] == ["?", "?", "?", "?", "?", 5, 0, "?"]
    assert D["y"] == [1, 2, 3, "?", 4, 5, 6, 7]
    D2, _ = io.read_sequences(str(p), masked_symbol="N", replace_mchar=None)
    assert D2["x"][:5] == ["-", "?", "-1", "N", "-3"]


def test_write_then_read_sequences(tmp_path):
    D = {"c1": [0, 3, "?"], "c2": [1, 0, 2]}
    p = tmp_path / "w.csv"
    io.write_sequences(D, 3, str(p))
    assert p.read_text().splitlines()[0] == "cell,r0,r1,r2"
    back, names = io.read_sequences(str(p))
    assert back == D and names == ["r0", "r1", "r2"]


@pytest.mark.parametrize("name", ["test4_prior.csv", "example1/priors.csv"])
def test_read_priors_csv_matches_the_fixture_reader(name):
    path = golden_path(name)
    Q = io.read_priors(path, {})
    want = read_priors_csv(path, len(Q))
    assert [{k: v for k, v in q.items() if k != 0} for q in want] == Q


def test_read_priors_csv_with_header_and_named_sites(tmp_path):
    p = tmp_path / "q.csv"
    p.write_text("site,state,prob\nr3,1,0.5\nr3,2,0.5\nr7,4,1.0\n")
    assert io.read_priors(str(p), {}) == [{1: 0.5, 2: 0.5}, {4: 1.0}]


def test_read_priors_from_dicts(tmp_path):
    priors = {0: {1: 0.25, 2: 0.75}, 1: {5: 1.0}}
    pj = tmp_path / "q.json"
    pj.write_text(json.dumps({str(k): {str(s): p for s, p in v.items()} for k, v in priors.items()}))
    assert io.read_priors(str(pj), {}) == [{1: 0.25, 2: 0.75, 0: 0}, {5: 1.0, 0: 0}]
    pk = tmp_path / "q.pkl"
    with open(pk, "wb") as f:
        pickle.dump(priors, f)
    # site names that do not match the keys but as many: sorted keys in order
    msa = {"a": [1, 5], "b": [2, 0]}
    assert io.read_priors(str(pk), msa, site_names=["r10", "r11"]) == [{1: 0.25, 2: 0.75, 0: 0}, {5: 1.0, 0: 0}]
    # more names than keys: by the digits of the names, uniform priors where a key is missing
    msa3 = {"a": [1, 5, 3], "b": [2, 0, 4]}
    Q = io.read_priors(str(pk), msa3, site_names=["r0", "r1", "r2"])
    assert Q[0] == {1: 0.25, 2: 0.75} and Q[1] == {5: 1.0} and Q[2] == {3: 0.5, 4: 0.5, 0: 0}


def test_dedup_and_add_dup():
    msa = {"a": [1, 0], "b": [1, 0], "c": [0, 2], "d": [1, 0], "e": [0, 3]}
    trees = ["(((a:1,b:1):1,c:2):1,(d:1,e:1):2);"]
    has, dup_map, _, pruned = io.dedup(msa, trees)
    assert has and dup_map == {"a": ["b", "d"]}
    t = read_tree_newick(pruned[0])
    assert sorted(v.label for v in t.traverse_leaves()) == ["a", "c", "e"]
    # b and d pruned, unifurcations spliced (lengths added)
    lens = {v.label: v.edge_length for v in t.traverse_leaves()}
    assert lens == {"a": 2, "c": 2, "e": 3}
    back = io.add_dup(pruned, dup_map)
    t2 = read_tree_newick(back[0])
    dummy = next(v for v in t2.traverse_postorder() if v.label == "node_dummy_a")
    assert sorted(c.label for c in dummy.children) == ["a", "b", "d"]
    assert dummy.edge_length == 2 and all(c.edge_length == 0.005 for c in dummy.children)
    assert io.dedup({"x": [1], "y": [2]}, trees)[0] is False


def test_tree_writers():
    # height 5 (the root's own branch counts) over timescale 2
    text, rate = io.scaled_trees_text(["((a:1,b:3):1.5,c:2):0.5;"], timescale=2.0)
    assert rate == 2.5
    assert text == "(((a:0.4,b:1.2):0.6,c:0.8):0.2);\n"
    # the 0.004 branch (<= 1.2 dmin) is contracted (its length dropped), then
    # everything is scaled by height 3.5
    col = io.collapsed_trees_text(["((a:1,b:3):0.004,c:2):0.5;"], timescale=1.0)
    t = read_tree_newick(col.strip()[1:-2] + ";")
    assert sorted(v.label for v in t.root.children) == ["a", "b", "c"]
    assert abs(t.height() - 1.0) < 1e-12
