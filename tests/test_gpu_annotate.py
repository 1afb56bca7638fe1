"""Ancestral annotation output (runner.py:367-435) from device posteriors
against the same output from the oracle's E-step: identical text, and the
per-node posteriors P(0 | D), P(-1 | D) to 1e-9."""
import math

import numpy as np
import pytest

from helpers import load_test1, load_test4
from laml_amd import sim
from laml_amd.annotate import annotation_text, node_values_device, write_annotations
from laml_amd.solvers import EM_solver
from oracle.em_oracle import OracleEM, postorder

pytestmark = pytest.mark.gpu


def _cases():
    tree, Q, codes, _ = sim.make_problem(120, 25, 3, 20, 1.0, 0.1, 0.2, 61)
    yield "syn120x25", tree.newick(), sim.codes_to_charmtrx(tree, codes), Q, 0.1, 0.2
    T, msa, Q = load_test4()
    yield "test4", T, msa, Q, 0.05, 0.1
    T, msa, Q = load_test1()
    yield "test1", T, msa, Q, 0.15877477685098262, 0.05231954386883335


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_annotation_text_matches_oracle(case):
    name, T, msa, Q, nu, phi = case
    s = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"nu": nu, "phi": phi})
    dev = node_values_device(s)
    text_dev = annotation_text(s.trees, dev, Q)
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    o.az_partition()
    o.estep()
    s2 = EM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"nu": nu, "phi": phi})
    vals = []
    worst = 0.0
    for t2, ro, dv in zip(s2.trees, o.trees, dev):
        d = {}
        dev_nodes = list(dv.values())
        for (v2, vo), (lab, p0, p1, S) in zip(zip(t2.traverse_postorder(), postorder(ro)), dev_nodes):
            d[v2] = (list(vo.a["alpha"]), vo.a["post0"], vo.a["post1"], [sum(vo.a["S"][j]) for j in range(5)])
            assert [str(x) for x in lab] == [str(x) for x in vo.a["alpha"]]
            for a, b in zip(p0, vo.a["post0"]):
                worst = max(worst, abs(math.exp(a) - math.exp(b)))
            for a, b in zip(p1, vo.a["post1"]):
                worst = max(worst, abs(math.exp(a) - math.exp(b)))
        vals.append(d)
    text_or = annotation_text(s2.trees, vals, Q)
    print(name, "max posterior difference", worst, flush=True)
    assert worst < 1e-9
    assert text_dev == text_or
    assert text_dev.count("\n") == 1 + sum(1 for _ in s.trees[0].traverse_preorder())


def test_write_annotations_file(tmp_path):
    tree, Q, codes, _ = sim.make_problem(30, 6, 3, 8, 1.0, 0.1, 0.2, 3)
    out = tmp_path / "ann.txt"
    text = write_annotations([tree.newick()], {"charMtrx": sim.codes_to_charmtrx(tree, codes)}, {"Q": Q},
                             {"nu": 0.1, "phi": 0.2}, str(out))
    assert out.read_text() == text and text.splitlines()[0].endswith(";")
