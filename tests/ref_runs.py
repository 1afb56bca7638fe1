"""Inputs and logged results of the reference's own example runs
(tests/golden/example1, example3, example2_api, reference_runs.json), shared
by the CPU and GPU tests that replay them."""
import json
import math

from helpers import golden_path
from laml_amd import io
from laml_amd.tree import read_tree_newick


def logged():
    with open(golden_path("reference_runs.json")) as f:
        return json.load(f)


def inputs(name):
    """(starting tree text, character matrix, priors) as the reference's
    runner reads them (comma-delimited matrix, '?' missing, CSV priors)."""
    msa, names = io.read_sequences(golden_path(name, "character_matrix.csv"), masked_symbol="?")
    Q = io.read_priors(golden_path(name, "priors.csv"), msa, site_names=names)
    T = open(golden_path(name, "starting.tree")).read().strip()
    return T, msa, Q


def params_file(name, prefix):
    out = {}
    with open(golden_path(name, prefix + "_params.txt")) as f:
        for line in f:
            k, v = line.split(":", 1)
            out[k.strip()] = v.strip()
    return out


def unscaled_trees(name, prefix, rate):
    """The optimised trees before the writer's scaling: each written tree
    (stem wrapped in an extra pair of parentheses) unwrapped and multiplied
    back by the mutation rate."""
    out = []
    with open(golden_path(name, prefix + "_trees.nwk")) as f:
        for line in f:
            if not line.strip():
                continue
            t = read_tree_newick(line.strip())
            r = t.root.children[0]
            r.parent = None
            t.root = r
            for v in t.traverse_preorder():
                v.edge_length = v.edge_length * rate
            out.append(t.newick())
    return out


def zero_fraction_dmax(msa):
    """LAML 0.0.2's dmax: -log of the fraction of zero entries."""
    rows = list(msa.values())
    zeros = sum(r.count(0) for r in rows)
    return -math.log(zeros / (len(rows) * len(rows[0])))
