"""Brute-force PMM likelihood and edge posteriors by enumerating the states of
every internal node (TEST INFRASTRUCTURE ONLY, tiny trees).

Independent of any inside/outside recursion: it sums the full joint
probability over all internal-state assignments, with the PMM transition
probabilities of laml_libs/sim_lib.py:72-92 (the generative model the
reference's EM_solver.py:91-300 marginalises) and the leaf emission model
(state s != -1 is observed as s w.p. 1-phi and '?' w.p. phi; state -1 is
observed as '?'; an explicit -1 entry means silenced).  The unedited
ancestor sits above the root's stem edge.
"""
import itertools
import math

from .em_oracle import parse_newick, postorder

SIL = "S"


def _trans(s, t, d, nu, q):
    p = math.exp(-d)
    if s == SIL:
        return 1.0 if t == SIL else 0.0
    if s == 0:
        if t == 0:
            return p ** (nu + 1)
        if t == SIL:
            return 1 - p ** nu
        return q.get(t, 0.0) * p ** nu * (1 - p)
    # edited
    if t == s:
        return p ** nu
    if t == SIL:
        return 1 - p ** nu
    return 0.0


def _emit(t, c, phi):
    if c == "?":
        return 1.0 if t == SIL else phi
    if c == -1:
        return 1.0 if t == SIL else 0.0
    if t == SIL:
        return 0.0
    return (1 - phi) if t == c else 0.0


def enumerate_site(tree_str, column, q, nu, phi):
    """column: dict leaf label -> observed entry at this site; q: edit-state
    priors (keys != 0).  Returns (L, counts, rtilde) with counts a dict
    node-label -> [S0..S4] and rtilde the sum over '?' leaves of
    P(leaf not silenced | D)."""
    root = parse_newick(tree_str)
    nodes = postorder(root)
    internal = [v for v in nodes if v.children]
    leaves = [v for v in nodes if not v.children]
    states = [0] + [a for a in q if a != 0] + [SIL]
    L = 0.0
    joint = {id(v): {} for v in nodes}  # (s_parent, s_v) -> prob
    leaf_sil = {id(v): 0.0 for v in leaves}
    for assign in itertools.product(states, repeat=len(internal)):
        st = {id(v): s for v, s in zip(internal, assign)}
        w = 1.0
        for v in internal:
            sp = 0 if v.parent is None else st[id(v.parent)]
            w *= _trans(sp, st[id(v)], v.edge_length, nu, q)
            if w == 0.0:
                break
        if w == 0.0:
            continue
        marg = {}
        for v in leaves:
            sp = st[id(v.parent)]
            m = 0.0
            for t in states:
                m += _trans(sp, t, v.edge_length, nu, q) * _emit(t, column[v.label], phi)
            marg[id(v)] = m
            w *= m
        if w == 0.0:
            continue
        L += w
        for v in internal:
            sp = 0 if v.parent is None else st[id(v.parent)]
            key = (sp, st[id(v)])
            joint[id(v)][key] = joint[id(v)].get(key, 0.0) + w
        for v in leaves:
            sp = st[id(v.parent)]
            for t in states:
                pt = _trans(sp, t, v.edge_length, nu, q) * _emit(t, column[v.label], phi)
                if pt == 0.0:
                    continue
                share = w * pt / marg[id(v)]
                joint[id(v)][(sp, t)] = joint[id(v)].get((sp, t), 0.0) + share
                if t != SIL and column[v.label] == "?":
                    leaf_sil[id(v)] += share
    if L == 0.0:
        return 0.0, None, None
    counts = {}
    for v in nodes:
        S = [0.0] * 5
        for (a, b), pr in joint[id(v)].items():
            pr /= L
            if a == 0 and b == 0:
                S[0] += pr
            elif a == 0 and b == SIL:
                S[2] += pr
            elif a == 0:
                S[1] += pr
            elif a != SIL and b == a:
                S[3] += pr
            elif a != SIL and b == SIL:
                S[4] += pr
        counts[v.label] = S
    rtilde = sum(x / L for x in leaf_sil.values())
    return L, counts, rtilde


def brute_nllh(tree_str, charMtrx, Q, nu, phi):
    K = len(Q)
    tot = 0.0
    for k in range(K):
        col = {c: row[k] for c, row in charMtrx.items()}
        s = sum(Q[k].values())
        q = {a: Q[k][a] / s for a in Q[k] if a != 0}
        L, _, _ = enumerate_site(tree_str, col, q, nu, phi)
        tot += math.log(L)
    return -tot
