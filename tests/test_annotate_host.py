"""The annotation writer's formatting rules (runner.py:368-388) on hand-worked
cases, and the text layout from given node values (CPU)."""
import math

from laml_amd.annotate import annotation_text, format_posterior, params_text
from laml_amd.tree import read_tree_newick


def test_format_posterior_rules():
    assert format_posterior(1, 0, 0, '3', {3: 1.0}) == '0'
    assert format_posterior(0, 1, 0, '3', {3: 1.0}) == '-1'
    assert format_posterior(0, 0, 1, '3', {3: 1.0}) == '3'
    assert format_posterior(0.25, 0.5, 0.25, '7', {7: 1.0}) == '0:0.25/-1:0.5/7:0.25'
    assert format_posterior(0.4, 0, 0.6, '?', {1: 0.5, 2: 0.4999, 9: 0.0001}) == '0:0.4/1:0.3/2:0.3'
    assert format_posterior(0, 0.3, 0.7, 'z', {1: 1.0}) == '-1:0.3/z:0.7'


def test_annotation_text_layout():
    t = read_tree_newick("((a:1,b:1):1,c:1):0.5;")
    nodes = list(t.traverse_postorder())
    lg = math.log
    vals = {}
    for v in nodes:
        lab = ['z'] if v.label in ('a', 'b') else ['?']
        vals[v] = (lab, [lg(0.5)], [lg(0.25)], [0.0, 0.1234, 0.2, 0.0, 0.3])
    text = annotation_text([t], [vals], [{1: 1.0}])
    lines = text.splitlines()
    # new root I0 above the root; internal nodes named I1, I2 in treeswift's
    # preorder (last child first: I0, I1, c, I2, b, a)
    assert lines[0] == "(((a:0.623,b:0.623)I2:0.623,c:0.623)I1:0.623)I0;"
    assert [x.split(",")[0] for x in lines[1:]] == ["I0", "I1", "c", "I2", "b", "a"]
    assert lines[1] == "I0,0"
    assert lines[2] == "I1,0:0.5/-1:0.25/1:0.25"
    assert lines[6] == "a,0:0.5/-1:0.25/z:0.25"
    assert params_text(0.1, 0.2, 3.5, 1.0).splitlines()[2] == "Negative-llh: 3.5"
