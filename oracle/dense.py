"""Dense per-state restatement of the PMM likelihood and E-step (TEST
INFRASTRUCTURE ONLY).

This is the Felsenstein form of the reference's fastEM path
(fast_em/calculations.py: leaf initialisation :278-334, inside :193-276,
outside :135-191, responsibilities :373-454), in FP64 and without its EPS
clamps: every (node, site) carries a vector over all states of the site
(unedited, each edit state of the site's alphabet, silenced), transitions are
full matrices, and posterior transition counts come from inside x outside
products.  It shares no code with the compressed restatement in
``em_oracle`` and makes no assumption about the priors summing to one, so it
pins both the compressed recursion (reference EM model) and the model with
unnormalised observed-state alphabets (fastEM model).

Sites are padded to the largest alphabet with zero-prior states, which are
unreachable.
"""
import math

import numpy as np

from .em_oracle import parse_newick, postorder, suppress_unifurcations


def site_tables(Q):
    """Per-site state lists and a padded prior matrix [K, M] (column j is
    edit state states[k][j])."""
    states = [[s for s in q if s != 0] for q in Q]
    M = max(1, max(len(s) for s in states))
    prior = np.zeros((len(Q), M))
    for k, q in enumerate(Q):
        for j, s in enumerate(states[k]):
            prior[k, j] = q[s]
    return states, prior


def transition(d, nu, prior):
    """P[k, s, t] over states (0, edits 1..M, S) for a branch of length d."""
    K, M = prior.shape
    n = M + 2
    P = np.zeros((K, n, n))
    e1 = math.exp(-d)
    en = math.exp(-nu * d)
    P[:, 0, 0] = e1 * en
    P[:, 0, 1:M + 1] = prior * (1.0 - e1) * en
    P[:, 0, M + 1] = 1.0 - en
    for j in range(M):
        P[:, j + 1, j + 1] = en
        P[:, j + 1, M + 1] = 1.0 - en
    P[:, M + 1, M + 1] = 1.0
    return P


def leaf_vector(row, states, M, phi, missing=("?",), silenced=(-1,)):
    """Emission P(observed | leaf state) [K, M+2]."""
    K = len(row)
    e = np.zeros((K, M + 2))
    for k, x in enumerate(row):
        if x in missing:
            e[k, :M + 1] = phi
            e[k, M + 1] = 1.0
        elif x in silenced:
            e[k, M + 1] = 1.0
        elif x == 0:
            e[k, 0] = 1.0 - phi
        else:
            e[k, 1 + states[k].index(x)] = 1.0 - phi
    return e


def dense_estep(tree_str, charMtrx, Q, nu, phi, missing=("?",), silenced=(-1,)):
    """NLL, per-branch counts S0..S4 summed over sites (post-order nodes) and
    the sum over missing leaf entries of P(not silenced | D)."""
    root = suppress_unifurcations(parse_newick(tree_str))
    nodes = postorder(root)
    states, prior = site_tables(Q)
    K, M = prior.shape
    n = M + 2
    P = {id(v): transition(v.edge_length, nu, prior) for v in nodes}
    beta = {}
    lsc = {}
    for v in nodes:
        if not v.children:
            b = leaf_vector(charMtrx[v.label], states, M, phi, missing, silenced)
            s = np.zeros(K)
        else:
            b = np.ones((K, n))
            s = np.zeros(K)
            for c in v.children:
                b = b * np.einsum("kst,kt->ks", P[id(c)], beta[id(c)])
                s = s + lsc[id(c)]
        m = b.max(axis=1)
        m[m == 0] = 1.0
        beta[id(v)] = b / m[:, None]
        lsc[id(v)] = s + np.log(m)
    top = np.einsum("kt,kt->k", P[id(root)][:, 0, :], beta[id(root)])
    logL = np.log(top) + lsc[id(root)]
    nll = -float(logL.sum())
    # outside: gamma[v][k, s] = P(data outside v's subtree, v = s) / scale
    gam = {}
    gsc = {}
    gam[id(root)] = P[id(root)][:, 0, :].copy()
    gsc[id(root)] = np.zeros(K)
    for v in reversed(nodes):
        for c in v.children:
            sib = [w for w in v.children if w is not c]
            g = gam[id(v)].copy()
            s = gsc[id(v)].copy()
            for w in sib:
                g = g * np.einsum("kst,kt->ks", P[id(w)], beta[id(w)])
                s = s + lsc[id(w)]
            gc = np.einsum("ks,kst->kt", g, P[id(c)])
            m = gc.max(axis=1)
            m[m == 0] = 1.0
            gam[id(c)] = gc / m[:, None]
            gsc[id(c)] = s + np.log(m)
            # the parent-side factor of the branch posterior
            c.a["pside"] = (g, s)
    counts = np.zeros((len(nodes), 5))
    rt = 0.0
    for i, v in enumerate(nodes):
        if v is root:
            g = np.zeros((K, n))
            g[:, 0] = 1.0
            s = np.zeros(K)
        else:
            g, s = v.a["pside"]
        # xi[k, a, b] = g[a] P[a, b] beta_v[b] / L
        w = np.exp(s + lsc[id(v)] - logL)
        xi = g[:, :, None] * P[id(v)] * beta[id(v)][:, None, :] * w[:, None, None]
        counts[i, 0] = xi[:, 0, 0].sum()
        counts[i, 1] = xi[:, 0, 1:M + 1].sum()
        counts[i, 2] = xi[:, 0, M + 1].sum()
        counts[i, 3] = np.einsum("kaa->", xi[:, 1:M + 1, 1:M + 1])
        counts[i, 4] = xi[:, 1:M + 1, M + 1].sum()
        if not v.children:
            row = charMtrx[v.label]
            post = gam[id(v)] * beta[id(v)] * np.exp(gsc[id(v)] + lsc[id(v)] - logL)[:, None]
            for k, x in enumerate(row):
                if x in missing:
                    rt += post[k, :M + 1].sum()
    return nll, counts, rt
