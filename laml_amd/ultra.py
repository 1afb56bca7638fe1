"""Ultrametric branch-length constraints.

``ultrametric_constr`` restates the reference's equality system
(laml_libs/EM_solver.py:53-89): for every internal node, the path lengths from
its two children down to a leaf (through first children) are equal; fixed
branches enter as constants, polytomy-resolution branches as zero; with several
trees the root path lengths (stem included) are equal.

``HeightModel`` is the device form of the same constraint set: node heights
above the leaves are the variables, so every ultrametric branch-length vector
is d_v = t(parent) - t(v).  Branches that are fixed or polytomy-marked tie two
heights (t(parent) = t(v) + c); they are contracted into *height classes*.
Classes containing a leaf are pinned (height known); the others are the
variables.  Each problem gets one extra class, the stem top, above its
roots, so that all roots' stem-inclusive heights are equal.  Free branches
join classes into a tree, which is what the device Newton solver eliminates
in two level sweeps.
"""
import numpy as np


def ultrametric_constr(trees, local_brlen_opt=True):
    """EM_solver.ultrametric_constr (EM_solver.py:53-89) for trees whose nodes
    carry ``polytomy_mark`` / ``mark_fixed``.  Returns (M, b) as lists."""
    N = 0
    for t in trees:
        for v in t.traverse_postorder():
            if not v.polytomy_mark and not (getattr(v, "mark_fixed", False) and local_brlen_opt):
                N += 1
    constrs = {}
    idx = 0
    path = {}
    const = {}
    for t in trees:
        for v in t.traverse_postorder():
            fixed = getattr(v, "mark_fixed", False)
            if not v.children:
                vec = [0.0] * N
                c0 = v.edge_length if fixed else 0
            else:
                c1, c2 = v.children
                m = tuple(x - y for x, y in zip(path[c1], path[c2]))
                mc = tuple(-x for x in m)
                c = const[c2] - const[c1]
                if sum(x != 0 for x in m) > 0 and not (m in constrs or mc in constrs):
                    constrs[m] = c
                vec = path[c1]
                c0 = const[c1] + v.edge_length if fixed else const[c1]
            if not v.polytomy_mark and not (fixed and local_brlen_opt):
                vec[idx] = 1
                idx += 1
            path[v] = vec
            const[v] = c0
    for t in trees[1:]:
        r0 = trees[0].root
        m = tuple(x - y for x, y in zip(path[r0], path[t.root]))
        constrs[m] = const[t.root] - const[r0]
    M = [list(m) for m in constrs]
    b = [constrs[m] for m in constrs]
    return M, b


class UltraInfeasible(Exception):
    pass


class HeightModel:
    """Height classes of one forest (all problems), built from the forest's
    level-ordered arrays.

    Attributes (numpy, level-order node ids):
      up_cls, up_off  class and offset of the parent side of each branch
                      (class -1: pinned, up_off is the absolute height)
      dn_cls, dn_off  the same for the child side
      cls_*           per variable class: problem, class level, parent class,
                      the branch to the parent class, incident free branches
    """

    def __init__(self, layout, fixed_len):
        """layout: ForestLayout; fixed_len: [N] branch length of every node
        (used for fixed branches; polytomy branches count as 0 and are passed
        with ``layout.poly``)."""
        N = layout.num_nodes
        par = layout.parent.astype(np.int64)
        free = layout.free.astype(bool)
        poly = getattr(layout, "poly", np.zeros(N, dtype=bool))
        P = layout.num_problems
        tree_prob = layout.tree_prob
        node_prob = tree_prob[layout.tree_of]
        # union-find over tied edges; node N + p is problem p's stem top
        tot = N + P
        uf = np.arange(tot)
        off = np.zeros(tot)  # t(x) = t(root_of_class) + off(x) after compression

        def find(x):
            path = []
            while uf[x] != x:
                path.append(x)
                x = uf[x]
            root = x
            # compress with offsets
            acc = 0.0
            for y in reversed(path):
                acc += off[y]
                off[y] = acc
                uf[y] = root
            return root

        def union(child, parent, c):
            # t(parent) = t(child) + c
            rc, rp = find(child), find(parent)
            oc, op = off[child] if child != rc else 0.0, off[parent] if parent != rp else 0.0
            if rc == rp:
                if abs((op - oc) - c) > 1e-12 * max(1.0, abs(c)):
                    raise UltraInfeasible("inconsistent fixed branch lengths")
                return
            # attach rc under rp: t(rc) = t(rp) + op - c - oc
            uf[rc] = rp
            off[rc] = op - c - oc

        up_node = np.where(par >= 0, par, N + node_prob)  # parent side node (stem top for roots)
        for v in range(N):
            if free[v]:
                continue
            c = 0.0 if poly[v] else float(fixed_len[v])
            union(v, int(up_node[v]), c)
        roots = np.array([find(x) for x in range(tot)])
        offs = np.array([off[x] if roots[x] != x else 0.0 for x in range(tot)])
        # pinned classes: those containing a leaf (t = 0)
        pinned = {}
        for v in range(layout.num_leaves):
            r = roots[v]
            h = -offs[v]  # t(root) such that t(v) = 0
            if r in pinned and abs(pinned[r] - h) > 1e-12 * max(1.0, abs(h)):
                raise UltraInfeasible("fixed branches pin a height class to two values")
            pinned[r] = h
        # variable classes: unpinned classes that some free branch touches
        touched = set()
        for v in range(N):
            if free[v]:
                touched.add(int(roots[int(up_node[v])]))
                touched.add(int(roots[v]))
        var_roots = sorted(touched - set(pinned))
        cls_id = {r: i for i, r in enumerate(var_roots)}
        V = len(var_roots)

        def side(x):
            r = roots[x]
            if r in pinned:
                return -1, pinned[r] + offs[x]
            return cls_id[r], offs[x]

        self.up_cls = np.full(N, -1, dtype=np.int32)
        self.up_off = np.zeros(N)
        self.dn_cls = np.full(N, -1, dtype=np.int32)
        self.dn_off = np.zeros(N)
        for v in range(N):
            if not free[v]:
                continue
            self.up_cls[v], self.up_off[v] = side(int(up_node[v]))
            self.dn_cls[v], self.dn_off[v] = side(v)
            if self.up_cls[v] == -1 and self.dn_cls[v] == -1:
                d = self.up_off[v] - self.dn_off[v]
                if not (layout.dmin <= d <= layout.dmax):
                    raise UltraInfeasible("a free branch between two pinned heights is out of bounds")
        # class tree: parent class through the unique free branch joining it upward
        cls_prob = np.array([node_prob[r] if r < N else r - N for r in var_roots], dtype=np.int32)
        cls_parent = np.full(V, -1, dtype=np.int32)
        cls_pedge = np.full(V, -1, dtype=np.int32)
        inc = [[] for _ in range(V)]
        for v in range(N):
            if not free[v]:
                continue
            u, d_ = self.up_cls[v], self.dn_cls[v]
            if u >= 0:
                inc[u].append((v, 1))
            if d_ >= 0:
                inc[d_].append((v, -1))
                if u >= 0:
                    if cls_parent[d_] != -1:
                        raise UltraInfeasible("height classes do not form a tree")
                    cls_parent[d_] = u
                    cls_pedge[d_] = v
        # class levels (children before parents)
        lvl = np.zeros(V, dtype=np.int64)
        order = []
        # topological: repeatedly take classes whose children are placed
        nchild = np.zeros(V, dtype=np.int64)
        for c in range(V):
            if cls_parent[c] >= 0:
                nchild[cls_parent[c]] += 1
        ready = [c for c in range(V) if nchild[c] == 0]
        while ready:
            nxt = []
            for c in ready:
                order.append(c)
                p = cls_parent[c]
                if p >= 0:
                    lvl[p] = max(lvl[p], lvl[c] + 1)
                    nchild[p] -= 1
                    if nchild[p] == 0:
                        nxt.append(p)
            ready = nxt
        if len(order) != V:
            raise UltraInfeasible("cycle in the class tree")
        # renumber classes by (level, problem, id)
        perm = np.lexsort((np.arange(V), cls_prob, lvl))
        newid = np.empty(V, dtype=np.int64)
        newid[perm] = np.arange(V)
        remap = lambda a: np.where(a >= 0, newid[np.maximum(a, 0)], -1).astype(np.int32)
        self.up_cls = remap(self.up_cls)
        self.dn_cls = remap(self.dn_cls)
        self.cls_parent = remap(cls_parent[perm])
        self.cls_pedge = cls_pedge[perm].astype(np.int32)
        self.cls_prob = cls_prob[perm]
        lv = lvl[perm]
        H = int(lv.max()) + 1 if V else 0
        self.cls_level_off = np.searchsorted(lv, np.arange(H + 1)).astype(np.int32)
        off_ = [0]
        ib, isg = [], []
        for c in perm:
            for (v, sg) in inc[c]:
                ib.append(v)
                isg.append(sg)
            off_.append(len(ib))
        self.inc_off = np.asarray(off_, dtype=np.int32)
        self.inc_branch = np.asarray(ib, dtype=np.int32)
        self.inc_sign = np.asarray(isg, dtype=np.int8)
        self.num_classes = V
        # per-problem class ranges are not contiguous; the device iterates all
        # classes of a level and skips other problems
        self.prob_has_free = np.zeros(P, dtype=np.uint8)
        for v in range(N):
            if free[v]:
                self.prob_has_free[node_prob[v]] = 1

    def lengths(self, t, N):
        """Branch lengths of the free branches for class heights t (others 0)."""
        up = np.where(self.up_cls >= 0, t[np.maximum(self.up_cls, 0)] + self.up_off, self.up_off)
        dn = np.where(self.dn_cls >= 0, t[np.maximum(self.dn_cls, 0)] + self.dn_off, self.dn_off)
        return up - dn

    def check_strict(self, t, free, dmin, dmax):
        d = self.lengths(t, len(free))[free.astype(bool)]
        return bool(np.all(d > dmin) and np.all(d < dmax))

    def heights_from_lengths(self, d, dmin, dmax, scale=1.0):
        """Heights placed bottom-up: each class at the largest height its
        child branches ask for, a child branch asking for
        clip(scale * d, 1.01 dmin, 0.99 dmax).  Not necessarily strictly
        feasible (an upper bound may fail); see ``start_heights``."""
        V = self.num_classes
        t = np.zeros(V)
        lo = np.full(V, -np.inf)
        for c in range(V):
            for i in range(self.inc_off[c], self.inc_off[c + 1]):
                v = self.inc_branch[i]
                if self.inc_sign[i] != 1:
                    continue
                dc = self.dn_cls[v]
                tdn = (t[dc] + self.dn_off[v]) if dc >= 0 else self.dn_off[v]
                want = min(max(scale * float(d[v]), 1.01 * dmin), 0.99 * dmax)
                lo[c] = max(lo[c], tdn + want - self.up_off[v])
            if not np.isfinite(lo[c]):
                lo[c] = 0.0
            t[c] = lo[c]
        return t

    def start_heights(self, d, free, dmin, dmax):
        """A strictly feasible start for the barrier solver: the bottom-up
        placement of ``heights_from_lengths`` with the wanted lengths scaled
        down until every free branch lies strictly inside (dmin, dmax) (long
        paths can push a branch to a pinned leaf beyond dmax).  The
        constrained optimum does not depend on the start.  Raises
        UltraInfeasible when even the shortest placement fails."""
        for scale in (1.0, 0.5, 0.25, 0.1, 0.05, 0.02, 0.01, 0.0):
            t = self.heights_from_lengths(d, dmin, dmax, scale)
            if self.check_strict(t, free, dmin, dmax):
                return t
        raise UltraInfeasible("no strictly feasible ultrametric start")
