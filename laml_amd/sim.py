"""Seeded synthetic workloads: Yule trees and PMM-simulated character matrices.

The character simulation follows the semantics of the reference's
``simulate_seqs`` (laml_libs/sim_lib.py:64-105): going down the tree, an
unedited site stays unedited with probability p^(nu+1), mutates to edit state a
with q_a p^nu (1-p), or is silenced with 1-p^nu (p = e^{-mu d}); an edited site
keeps its state with p^nu or is silenced; silencing is absorbing; a leaf that
is not silenced drops out with probability phi.  Silenced and dropped entries
are observed as '?'.  Random numbers come from numpy's PCG64, vectorised over
the nodes of one depth and all sites (the reference draws with Python's
``random.choices`` one entry at a time), so the streams differ but the model is
the same.
"""
import numpy as np

from .forest import ProblemSpec, SiteTable, TreeArrays
from .tree import Node, Tree


class YuleTree:
    """Binary tree as parent/children arrays in post-order (root last)."""

    def __init__(self, c0, c1, parent, length, leaf_names):
        self.c0 = c0
        self.c1 = c1
        self.parent = parent
        self.length = length
        self.leaf_names = leaf_names  # post-order leaf index -> name

    @property
    def n_nodes(self):
        return len(self.c0)

    def to_tree(self, internal_labels=False):
        nodes = [Node() for _ in range(self.n_nodes)]
        li = 0
        ii = 0
        for v in range(self.n_nodes):
            nodes[v].edge_length = float(self.length[v])
            if self.c0[v] < 0:
                nodes[v].label = self.leaf_names[li]
                li += 1
            else:
                nodes[v].add_child(nodes[self.c0[v]])
                nodes[v].add_child(nodes[self.c1[v]])
                if internal_labels:
                    nodes[v].label = "I%d" % ii
                    ii += 1
        return Tree(nodes[-1])

    def newick(self, internal_labels=False):
        return self.to_tree(internal_labels).newick()


def yule_tree(n, rng, height=1.0):
    """Pure-birth (rate 1) tree grown to n leaves with a stem edge, scaled so
    that stem + root-to-tip = ``height`` (ultrametric)."""
    if n < 2:
        raise ValueError("need at least two leaves")
    # birth order: node ids; lineage k alive holds a node id whose branch grows
    max_nodes = 2 * n - 1
    parent = np.full(max_nodes, -1, dtype=np.int64)
    c0 = np.full(max_nodes, -1, dtype=np.int64)
    c1 = np.full(max_nodes, -1, dtype=np.int64)
    start = np.zeros(max_nodes)
    end = np.zeros(max_nodes)
    alive = [0]
    nxt = 1
    t = 0.0
    waits = rng.exponential(1.0, size=n)  # scaled by 1/k below
    picks = rng.random(n)
    while len(alive) < n:
        k = len(alive)
        t += waits[k - 1] / k
        j = int(picks[k - 1] * k)
        v = alive[j]
        end[v] = t
        a, b = nxt, nxt + 1
        nxt += 2
        parent[a] = v
        parent[b] = v
        c0[v] = a
        c1[v] = b
        start[a] = t
        start[b] = t
        alive[j] = a
        alive.append(b)
    t += waits[n - 1] / n
    for v in alive:
        end[v] = t
    length = (end - start)[:nxt] / t * height
    # post-order renumbering (children before parents, root last)
    order = []
    stack = [(0, False)]
    while stack:
        v, done = stack.pop()
        if done:
            order.append(v)
        else:
            stack.append((v, True))
            if c0[v] >= 0:
                stack.append((c1[v], False))
                stack.append((c0[v], False))
    order = np.asarray(order, dtype=np.int64)
    new = np.empty(nxt, dtype=np.int64)
    new[order] = np.arange(nxt)
    pc0 = np.where(c0[order] >= 0, new[np.maximum(c0[order], 0)], -1)
    pc1 = np.where(c1[order] >= 0, new[np.maximum(c1[order], 0)], -1)
    ppar = np.where(parent[order] >= 0, new[np.maximum(parent[order], 0)], -1)
    names = ["c%d" % i for i in range(n)]
    return YuleTree(pc0, pc1, ppar, length[order], names)


def depth_levels(tree):
    """Node ids grouped by depth from the root (root first)."""
    N = tree.n_nodes
    depth = np.zeros(N, dtype=np.int64)
    for v in range(N - 1, -1, -1):  # reverse post-order visits parents first
        p = tree.parent[v]
        if p >= 0:
            depth[v] = depth[p] + 1
    order = np.argsort(depth, kind="stable")
    bounds = np.searchsorted(depth[order], np.arange(depth.max() + 2))
    return [order[bounds[i]:bounds[i + 1]] for i in range(len(bounds) - 1)]


def random_priors(K, m_lo, m_hi, alpha, rng):
    """Per-site priors: alphabet size uniform in [m_lo, m_hi], q ~ Dirichlet(alpha)."""
    Q = []
    for _ in range(K):
        m = int(rng.integers(m_lo, m_hi + 1))
        q = rng.dirichlet(np.full(m, float(alpha)))
        Q.append({j + 1: float(q[j]) for j in range(m)})
    return Q


def simulate_codes(tree, Q, nu, phi, rng, mu=1.0):
    """PMM forward simulation; returns int16 codes [n_leaves, K] in leaf
    post-order with 0 unedited, j = edit state j, -1 '?' (silenced or dropped)."""
    K = len(Q)
    N = tree.n_nodes
    mmax = max(len(q) for q in Q)
    cum = np.ones((K, mmax))
    for k, q in enumerate(Q):
        c = np.cumsum([q[j + 1] for j in range(len(q))])
        cum[k, :len(q)] = c / c[-1]
    msize = np.asarray([len(q) for q in Q], dtype=np.int64)
    state = np.zeros((N, K), dtype=np.int32)  # -1 = silenced
    for lvl in depth_levels(tree):
        p_st = np.zeros((len(lvl), K), dtype=np.int32)
        par = tree.parent[lvl]
        has = par >= 0
        p_st[has] = state[par[has]]
        d = tree.length[lvl][:, None] * mu
        p = np.exp(-d)
        pn = p ** nu
        u = rng.random((len(lvl), K))
        u2 = rng.random((len(lvl), K))
        st = p_st.copy()
        zero = p_st == 0
        stay0 = p ** (nu + 1)
        mutp = pn * (1 - p)
        # unedited parent: stay / mutate / silence
        mut = zero & (u >= stay0) & (u < stay0 + mutp)
        sil0 = zero & (u >= stay0 + mutp)
        ri, ki = np.nonzero(mut)
        if len(ri):
            ns = (u2[ri, ki][:, None] > cum[ki]).sum(-1) + 1
            st[ri, ki] = np.minimum(ns, msize[ki])
        st = np.where(sil0, -1, st)
        # edited parent: keep or silence
        ed = p_st > 0
        st = np.where(ed & (u >= np.broadcast_to(pn, u.shape)), -1, st)
        state[lvl] = st
    leaves = np.nonzero(tree.c0 < 0)[0]
    obs = state[leaves].copy()
    drop = (obs >= 0) & (rng.random(obs.shape) < phi)
    obs[drop] = -1
    obs[obs < 0] = -1
    return obs.astype(np.int16)


def codes_to_charmtrx(tree, codes):
    """Reference-style character matrix dict (cell -> list with '?')."""
    out = {}
    for i, name in enumerate(tree.leaf_names):
        out[name] = [("?" if c < 0 else int(c)) for c in codes[i]]
    return out


def make_problem(n, K, m_lo, m_hi, alpha, nu, phi, seed, start_len=None):
    """A seeded single-tree workload.  Returns (YuleTree, Q, codes, truth)
    where truth holds the generating branch lengths, nu and phi."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tree = yule_tree(n, rng)
    Q = random_priors(K, m_lo, m_hi, alpha, rng)
    codes = simulate_codes(tree, Q, nu, phi, rng)
    truth = {"d": tree.length.copy(), "nu": nu, "phi": phi}
    return tree, Q, codes, truth


def problem_spec(tree, Q, codes, lengths=None, free=None, flags=3, sites=None):
    """ProblemSpec for the device forest straight from arrays (``sites``: a
    SiteTable of Q to share between problems with the same priors)."""
    sites = SiteTable(Q) if sites is None else sites
    L = tree.length if lengths is None else lengths
    F = np.ones(tree.n_nodes, dtype=np.uint8) if free is None else free
    ta = TreeArrays(tree.c0, tree.c1, tree.parent, L, F, codes)
    nobs = float((codes != -1).sum())
    phimin = 1.0 / (len(Q) * codes.shape[0])
    return ProblemSpec([ta], sites, nobs, phimin, flags)


def nni_variant(tree, codes, rng):
    """The tree after one random NNI move: a child c of a random internal
    non-root node u is swapped with u's sibling w.  Works on the arrays: the
    nodes keep their indices, except that u moves to just before its parent
    when w came after it (children must precede parents), so the leaves keep
    their order and the code rows need no permutation.  Returns (YuleTree,
    codes)."""
    c0, c1, par = tree.c0.copy(), tree.c1.copy(), tree.parent.copy()
    internal = np.nonzero((c0 >= 0) & (par >= 0))[0]
    if len(internal) == 0:
        raise ValueError("tree has no internal non-root branch")
    u = int(internal[int(rng.integers(len(internal)))])
    v = int(par[u])
    w = int(c1[v] if c0[v] == u else c0[v])
    if int(rng.integers(2)) == 0:
        c, c0[u] = int(c0[u]), w
    else:
        c, c1[u] = int(c1[u]), w
    if c0[v] == w:
        c0[v] = c
    else:
        c1[v] = c
    par[w] = u
    par[c] = v
    length = tree.length
    if w > u:
        # order: ... u-1, u+1 .. v-1, u, v ...
        N = len(c0)
        order = np.concatenate([np.arange(u), np.arange(u + 1, v), [u], np.arange(v, N)])
        new = np.empty(N, dtype=np.int64)
        new[order] = np.arange(N)

        def remap(a):
            a = a[order]
            return np.where(a >= 0, new[np.maximum(a, 0)], -1)

        c0, c1, par = remap(c0), remap(c1), remap(par)
        length = length[order]
    return YuleTree(c0, c1, par, length.copy(), list(tree.leaf_names)), codes


def scoring_workload(base_n, variants, others, n_lo, n_hi, K, m_lo, m_hi, alpha, nu, phi, seed):
    """Batched-scoring workload: `variants` NNI-perturbed copies of one seeded
    `base_n`-leaf Yule tree (sharing its characters and priors) followed by
    `others` independent Yule trees with log-uniform leaf counts in
    [n_lo, n_hi], each with its own priors and characters, all with K sites.
    Returns a list of (YuleTree, Q, codes)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    base_tree, Q, codes, _ = make_problem(base_n, K, m_lo, m_hi, alpha, nu, phi, seed)
    problems = []
    for _ in range(variants):
        t, cd = nni_variant(base_tree, codes, rng)
        problems.append((t, Q, cd))
    sizes = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), others)).astype(int)
    for i, n in enumerate(sizes):
        t, q, cd, _ = make_problem(int(n), K, m_lo, m_hi, alpha, nu, phi, seed + 1 + i)
        problems.append((t, q, cd))
    return problems
