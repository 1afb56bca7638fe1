"""Lowest common ancestors of leaf-label groups (integer work, bit-exact).

Replaces laml_libs/lca_lib.py:13-94 (``find_LCAs``): for every query, a tuple
of leaf labels, return the node that is the LCA of the labels present in the
tree (labels absent from the tree are ignored with a warning; a query with no
label in the tree maps to None).  The LCA is unique, so any exact algorithm
returns the reference's node; this one climbs parent pointers by depth.
"""
import logging

logger = logging.getLogger("laml_amd.lca")


def find_LCAs(tree, queries):
    by_label = {}
    depth = {}
    for v in tree.traverse_preorder():
        depth[v] = 0 if v.parent is None else depth[v.parent] + 1
        if v.label is not None:
            by_label[v.label] = v
    out = []
    for q in queries:
        nodes = []
        for a in q:
            v = by_label.get(a)
            if v is None:
                logger.warning("ignored calibration for taxon %s which is not found in the input tree", a)
            else:
                nodes.append(v)
        if not nodes:
            logger.warning("failed to find lca for %s", str(q))
            out.append(None)
            continue
        x = nodes[0]
        for y in nodes[1:]:
            a, b = x, y
            while depth[a] > depth[b]:
                a = a.parent
            while depth[b] > depth[a]:
                b = b.parent
            while a is not b:
                a = a.parent
                b = b.parent
            x = a
        out.append(x)
    return out
