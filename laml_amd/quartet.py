"""Quartet-local scoring of NNI candidates.

With ``local_brlen_opt`` (Topology_search.py:247-287) a candidate keeps nu,
phi and every branch except the five around the move at their current values
(u's children, v's children, v).  Its EM then changes nothing outside the
quartet: the subtrees hanging off it keep their inside vectors and the data
above v keeps its outside state.  ``score_moves`` therefore builds one device
forest for the tree *before* the move, runs one E-step on it, and lets
``laml_quartet_em`` run every candidate's EM on the resident values -- the
same EM (same start, same steps, same convergence test) as the
per-candidate solver, at O(K) per EM step instead of a pass over the tree.

Lengths follow the reference's candidate set-up: a fixed branch gets the
length its anchor pair names in the current tree text (``fixed_brlen``), a
free branch starts from the search tree's own length, and every branch is
raised to at least 2 dmin by ``ini_brlens`` (ML_solver.py:195-205).
"""
import numpy as np
import torch

from .lca import find_LCAs
from .tree import read_tree_newick

_BRANCHES = ("A", "W", "u", "C", "v")


def _clade_index(tree):
    """frozenset of leaf labels below each node -> post-order index."""
    out = {}
    below = {}
    for i, v in enumerate(tree.traverse_postorder()):
        if v.is_leaf():
            s = frozenset([v.label])
        else:
            s = frozenset().union(*[below[id(c)] for c in v.children])
        below[id(v)] = s
        out[s] = i
    return out, below


class QuartetResult:
    """What Topology_search reads from a scored candidate's solver."""

    def __init__(self, newicks, params):
        self._newicks = newicks
        self._params = params

    def get_tree_newick(self):
        return list(self._newicks)

    def get_params(self):
        return dict(self._params)


class QuartetScorer:
    """Scores NNI moves of a Topology_search's current trees (one problem)."""

    def __init__(self, search, strategy, solver_cls):
        self.search = search
        self.solver_cls = solver_cls
        self.nu = float(search.params['nu'])
        self.phi = float(search.params['phi'])
        trees = search.treeList_obj
        # anchors of every node of the current search trees (as _local_strategy)
        for t in trees:
            for x in t.traverse_postorder():
                if x.is_leaf():
                    x.anchors = (x.label, x.label)
                else:
                    C = x.child_nodes()
                    x.anchors = (C[0].anchors[0], C[-1].anchors[0])
        # fixed length of every node: its anchors' LCA in the current tree text
        self.base_len = []
        for ti, t in enumerate(trees):
            cur = read_tree_newick(search.treeTopoList[ti])
            nodes = list(t.traverse_postorder())
            lcas = find_LCAs(cur, [x.anchors for x in nodes])
            self.base_len.append({id(x): y.edge_length for x, y in zip(nodes, lcas)})
        # the base solver: the search trees as they are, fixed lengths, fixed nu/phi
        newicks = []
        saved = {}
        for ti, t in enumerate(trees):
            for x in t.traverse_postorder():
                saved[id(x)] = x.edge_length
                x.edge_length = self.base_len[ti][id(x)]
            newicks.append(t.newick())
            for x in t.traverse_postorder():
                x.edge_length = saved[id(x)]
        s0 = solver_cls(newicks, search.data, search.prior, search.params)
        self.dmin = s0.dmin
        for t in s0.trees:
            for v in t.traverse_postorder():
                v.mark_fixed = True
                if v.edge_length is not None and not v.polytomy_mark:
                    v.edge_length = max(2 * s0.dmin, v.edge_length)
        self.s0 = s0
        F = s0._forest_for(0)
        d, nu, phi = s0._device_params(F)
        nu.fill_(self.nu)
        phi.fill_(self.phi)
        # quartet EM reads every node's rows: no cherry fusion in this E-step
        F.set_fusion(False)
        F.nll(d, nu, phi, site_llh=True)
        sl = F.site_llh.cpu().numpy()
        self.tree_llh = [float(sl[F.site_base[i]:F.site_base[i + 1]].sum()) for i in range(F.num_trees)]
        F.estep(d, nu, phi)  # leave the outside state resident
        self.F, self.d, self.nu_t, self.phi_t = F, d, nu, phi
        self.clades = [_clade_index(t)[0] for t in s0.trees]
        self.search_below = [_clade_index(t)[1] for t in trees]

    def _gid(self, ti, clade):
        F = self.F
        return int(F.inv[F.tree_base[ti] + self.clades[ti][clade]])

    def score(self, moves, max_iter=1000, tol=1e-8):
        """moves: (u, v, u_child, w) of the search trees (before the move).
        Returns [(score, QuartetResult)] in the same order."""
        trees = self.search.treeList_obj
        tree_of = {}
        for ti, t in enumerate(trees):
            for x in t.traverse_postorder():
                tree_of[id(x)] = ti
        n = len(moves)
        q = np.zeros((n, 5), dtype=np.int32)
        L0 = np.zeros((n, 5))
        opt = np.full(n, 31, dtype=np.uint8)
        const = np.zeros(n)
        total = sum(self.tree_llh)
        for i, (u, v, uc, w) in enumerate(moves):
            ti = tree_of[id(u)]
            below = self.search_below[ti]
            A = next(x for x in u.child_nodes() if x is not uc)
            if v.parent is None:
                q[i, 0] = q[i, 1] = -1
            else:
                p = v.parent
                s = next(x for x in p.child_nodes() if x is not v)
                q[i, 0] = self._gid(ti, below[id(p)])
                q[i, 1] = self._gid(ti, below[id(s)])
            q[i, 2] = self._gid(ti, below[id(A)])
            q[i, 3] = self._gid(ti, below[id(w)])
            q[i, 4] = self._gid(ti, below[id(uc)])
            for j, x in enumerate((A, w, u, uc, v)):
                el = x.edge_length if x.edge_length is not None else 0.0
                L0[i, j] = max(2 * self.dmin, el)
            const[i] = total - self.tree_llh[ti]
        lo, nll, it = self.F.quartet_em(q, L0, opt, const, self.nu_t, self.phi_t, max_iter, tol)
        lo = lo.cpu().numpy()
        nll = nll.cpu().numpy()
        out = []
        for i, mv in enumerate(moves):
            out.append((-float(nll[i]), _LazyResult(self, mv, lo[i], tree_of[id(mv[0])])))
        self.iters = it.cpu().numpy()
        return out

    def close(self):
        if self.F is not None:
            self.F.close()
            self.F = None
            self.s0._forest = None


class _LazyResult:
    """Builds the candidate's tree text only when the search accepts it: the
    moved topology, fixed branches at their (raised) fixed lengths, the five
    free branches at their optimised lengths, branches of at most dmin/100
    collapsed (ML_solver.py:317-318)."""

    def __init__(self, scorer, move, lengths, ti):
        self.scorer = scorer
        self.move = move
        self.lengths = lengths
        self.ti = ti
        self.params = {'nu': scorer.nu, 'phi': scorer.phi}
        self.snapshot = None

    def materialise(self):
        """Tree texts of the candidate (call while the search trees still
        hold the pre-move topology)."""
        sc = self.scorer
        u, v, uc, w = self.move
        A = next(x for x in u.child_nodes() if x is not uc)
        free = {id(A): 0, id(w): 1, id(u): 2, id(uc): 3, id(v): 4}
        from .topology_search import _nni_swap, _nni_unswap
        _nni_swap(u, v, uc, w)
        try:
            newicks = []
            for ti, t in enumerate(sc.search.treeList_obj):
                saved = {}
                for x in t.traverse_postorder():
                    saved[id(x)] = x.edge_length
                    if ti == self.ti and id(x) in free:
                        x.edge_length = float(self.lengths[free[id(x)]])
                    else:
                        b = sc.base_len[ti][id(x)]
                        x.edge_length = max(2 * sc.dmin, b) if b is not None else None
                tc = read_tree_newick(t.newick())
                tc.collapse_short_branches(sc.dmin * 0.01)
                newicks.append(tc.newick())
                for x in t.traverse_postorder():
                    x.edge_length = saved[id(x)]
        finally:
            _nni_unswap(u, v, uc, w)
        return QuartetResult(newicks, self.params)
