"""Ancestral annotation output (laml_libs/runner.py:367-435).

After the optimisation, LAML re-runs the E-step on the optimal trees and
writes, per tree, the tree with each branch's length replaced by its expected
number of edit/silencing events (S1 + S2 + S4 summed over sites, rounded to
3 places), then one line per node (preorder) with its posterior state at
every site.  The E-step, the posteriors and the az-partition labels come from
the device; this module only formats them, with the reference's rules:

  - a tree whose root has more than one child gets a new root 'I0' above it
    (labels 'z', posteriors P(0) = 1);
  - unlabelled or repeated node labels become 'I<k>';
  - a '?' label inherits the parent's label when the parent's is not 'z';
  - P(0), P(-1) rounded to 2 places, P(edit) = round(1 - P(0) - P(-1), 2), and
    ``format_posterior`` writes the string.
"""
import math

import numpy as np

from .solvers import EM_solver
from .tree import Node


def format_posterior(p0, p_minus_1, p_alpha, alpha, q):
    """runner.py:368-388: a certain state is written alone, otherwise the
    non-zero parts 'state:prob' joined by '/'; for a '?' label the edit
    probability is spread over the site's prior."""
    if p0 == 1:
        return '0'
    elif p_minus_1 == 1:
        return '-1'
    elif p_alpha == 1:
        return alpha
    out = ''
    if p0 > 0:
        out += '0:' + str(p0)
    if p_minus_1 > 0:
        if out != '':
            out += '/'
        out += '-1:' + str(p_minus_1)
    if p_alpha > 0:
        if out != '':
            out += '/'
        if alpha == '?':
            out += "/".join([str(y) + ':' + str(round(p_alpha * q[y], 3)) for y in q if round(p_alpha * q[y], 3) > 0])
        else:
            out += alpha + ":" + str(p_alpha)
    return out


def node_values_device(solver):
    """Per tree: {node: (labels, post0 list, post1 list (log), S-sum list)}
    from one device E-step of ``solver`` (an EM_solver at its parameters)."""
    # the posteriors read every node's rows: no cherry fusion in this E-step
    solver._forest_for().set_fusion(False)
    solver.Estep()
    F = solver._forest
    d, nu, phi = solver._last_estep
    (l0, i0), (lS, iS) = F.posteriors(phi)
    labels = solver.node_labels()
    K = solver.numsites
    nl = F.num_leaves
    out = []
    for ti, t in enumerate(solver.trees):
        res = {}
        gids = solver._node_gids(F, ti)
        for pi, v in enumerate(t.traverse_postorder()):
            g = gids[pi]
            if g < nl:
                p0, pS = l0[g * K:(g + 1) * K], lS[g * K:(g + 1) * K]
            else:
                p0, pS = i0[(g - nl) * K:(g - nl + 1) * K], iS[(g - nl) * K:(g - nl + 1) * K]
            with np.errstate(divide="ignore"):
                res[v] = (labels[ti][v], list(np.log(p0)), list(np.log(pS)), list(v.S_sum))
        out.append(res)
    return out


def annotation_text(trees, values, Q):
    """The annotation file text (runner.py:389-428) for ``trees`` with per-node
    ``values`` (labels, post0, post1 as logs, S sums)."""
    lines = []
    idx = 0
    for tree, vals in zip(trees, values):
        K = len(Q)
        info = {}
        for v, (lab, p0, p1, S) in vals.items():
            info[v] = [list(lab), list(p0), list(p1), S]
        if len(tree.root.children) > 1:
            root = Node()
            root.label = 'I0'
            info[root] = [['z'] * K, [0.0] * K, [-float("inf")] * K, None]
            idx = 1
            root.add_child(tree.root)
            tree.root = root
        all_labels = set()
        for node in tree.traverse_preorder():
            if node.is_root():
                continue
            if node.label is None or node.label in all_labels:
                node.label = 'I' + str(idx)
                idx += 1
            all_labels.add(node.label)
            S = info[node][3]
            node.edge_length = round(S[1] + S[2] + S[4], 3)
        lines.append(tree.newick())
        for node in tree.traverse_preorder():
            lab, p0s, p1s, _ = info[node]
            post = ''
            for j in range(len(lab)):
                if not node.is_root():
                    plab = info[node.parent][0]
                    if lab[j] == '?' and plab[j] != 'z':
                        lab[j] = plab[j]
                p0 = round(math.exp(p0s[j]), 2)
                p_minus_1 = round(math.exp(p1s[j]), 2)
                p_alpha = round(1 - p0 - p_minus_1, 2)
                if post != '':
                    post += ','
                post += format_posterior(p0, p_minus_1, p_alpha, str(lab[j]), Q[j])
            lines.append(node.label + "," + str(post))
    return "\n".join(lines) + "\n"


def write_annotations(trees_newick, data, prior, params, out_path=None):
    """runner.py:389-428: annotate the optimal trees at the optimal
    parameters; returns the text and writes it to ``out_path`` if given."""
    s = EM_solver(trees_newick, data, prior, params)
    values = node_values_device(s)
    text = annotation_text(s.trees, values, prior['Q'])
    if out_path is not None:
        with open(out_path, "w") as f:
            f.write(text)
    return text


def params_text(phi, nu, nllh, mutation_rate):
    """The parameters file (runner.py:431-435)."""
    return ("Dropout rate: " + str(phi) + "\n" + "Silencing rate: " + str(nu) + "\n" +
            "Negative-llh: " + str(nllh) + "\n" + "Mutation rate: " + str(mutation_rate) + "\n")
