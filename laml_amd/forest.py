"""Host-side preparation of the device forest (integer work, bit-exact).

A *problem* is what one reference solver object holds: one or more trees that
share nu, phi and a character matrix (ML_solver.py:19-86).  A *forest* packs
any number of problems for one set of device launches.  This module

  - numbers the nodes of every tree in the reference's post-order (the order
    of the branch-length vector, EM_solver.py:45-51 / 328-348);
  - assigns height levels (longest path down to a leaf) and sorts all nodes of
    the forest by (height, tree, post-order index), so each level is one
    contiguous range and leaves come first;
  - recodes each site's edit states to 1..m (CSR-packed priors) and each leaf
    entry to an int16 code (0 unedited, j edit, -1 '?', -2 explicit -1);
  - builds the C-ABI description and creates the device forest.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import CODE_MISSING, CODE_SILENCED, LAML_OPT_NU, LAML_OPT_PHI

DMIN = 0.005  # laml_libs/core_constants.py:6
DMAX = 10.0   # laml_libs/core_constants.py:7


class TreeArrays:
    """Post-order arrays of one binary tree."""

    __slots__ = ("n_nodes", "n_leaves", "c0", "c1", "parent", "height", "length", "free",
                 "leaf_post", "codes", "nodes", "poly")

    def __init__(self, c0, c1, parent, length, free, codes, nodes=None, poly=None):
        self.c0 = np.asarray(c0, dtype=np.int64)
        self.c1 = np.asarray(c1, dtype=np.int64)
        self.parent = np.asarray(parent, dtype=np.int64)
        self.length = np.asarray(length, dtype=np.float64)
        self.free = np.asarray(free, dtype=np.uint8)
        self.n_nodes = len(self.c0)
        leaf = self.c0 < 0
        self.leaf_post = np.nonzero(leaf)[0]
        self.n_leaves = len(self.leaf_post)
        self.codes = codes  # int16 [n_leaves, K] rows in leaf post-order
        self.nodes = nodes
        self.poly = np.zeros(self.n_nodes, dtype=bool) if poly is None else np.asarray(poly, dtype=bool)
        # height of every node (edges on its longest path down to a leaf):
        # children precede parents, so sweeps of h[v] = 1 + max(h[children])
        # settle one more level each, bottom up
        h = np.zeros(self.n_nodes, dtype=np.int64)
        inner = np.nonzero(~leaf)[0]
        a, b = self.c0[inner], self.c1[inner]
        while True:
            hn = 1 + np.maximum(h[a], h[b])
            if np.array_equal(hn, h[inner]):
                break
            h[inner] = hn
        self.height = h


def tree_arrays_from_tree(tree, codes_of_leaf, free_of_node=None, poly_of_node=None):
    """Post-order arrays of a binary ``laml_amd.tree.Tree``; ``codes_of_leaf``
    maps a leaf Node to its int16 code row."""
    nodes = list(tree.traverse_postorder())
    idx = {id(v): i for i, v in enumerate(nodes)}
    N = len(nodes)
    c0 = np.full(N, -1, dtype=np.int64)
    c1 = np.full(N, -1, dtype=np.int64)
    parent = np.full(N, -1, dtype=np.int64)
    length = np.zeros(N, dtype=np.float64)
    free = np.ones(N, dtype=np.uint8)
    poly = np.zeros(N, dtype=bool)
    rows = []
    for i, v in enumerate(nodes):
        if poly_of_node is not None:
            poly[i] = bool(poly_of_node(v))
        if v.children:
            if len(v.children) != 2:
                raise ValueError("device forest requires binary trees (resolve polytomies first)")
            c0[i] = idx[id(v.children[0])]
            c1[i] = idx[id(v.children[1])]
        else:
            rows.append(codes_of_leaf(v))
        if v.parent is not None:
            parent[i] = idx[id(v.parent)]
        length[i] = 0.0 if v.edge_length is None else float(v.edge_length)
        if free_of_node is not None:
            free[i] = 1 if free_of_node(v) else 0
    if N < 3:
        raise ValueError("device forest requires trees with at least two leaves")
    codes = np.stack(rows).astype(np.int16) if rows else np.zeros((0, 0), np.int16)
    return TreeArrays(c0, c1, parent, length, free, codes, nodes, poly)


def resolved_tree_arrays(tree, codes_of_leaf, free_of_node=None):
    """Post-order arrays of a tree that may have polytomies, each resolved
    for the device into a balanced binary tree of zero-length fixed
    branches.  A zero-length branch has the identity transition (p = 1: no
    edit, no silencing), so an added node's inside values are the products
    of its children's, and the likelihood is that of the polytomy, whose
    parent multiplies over all its children (ML_solver.py:174-179).
    Returns (arrays, index of each node of ``tree`` in post-order)."""
    c0l, c1l, lengths, frees, polys, rows = [], [], [], [], [], []

    def add(a, b, length, free, poly):
        c0l.append(a)
        c1l.append(b)
        lengths.append(length)
        frees.append(free)
        polys.append(poly)
        return len(c0l) - 1

    def join(ids):
        if len(ids) == 1:
            return ids[0]
        h = len(ids) // 2
        return add(join(ids[:h]), join(ids[h:]), 0.0, 0, True)

    index = {}
    post_index = []
    for v in tree.traverse_postorder():
        length = 0.0 if v.edge_length is None else float(v.edge_length)
        free = 1 if free_of_node is None or free_of_node(v) else 0
        if not v.children:
            i = add(-1, -1, length, free, False)
            rows.append(codes_of_leaf(v))
        else:
            ch = [index[id(c)] for c in v.children]
            if len(ch) < 2:
                raise ValueError("device forest requires internal nodes with at least two children")
            h = len(ch) // 2
            i = add(join(ch[:h]), join(ch[h:]), length, free, False)
        index[id(v)] = i
        post_index.append(i)
    N = len(c0l)
    if N < 3:
        raise ValueError("device forest requires trees with at least two leaves")
    c0 = np.asarray(c0l, dtype=np.int64)
    c1 = np.asarray(c1l, dtype=np.int64)
    parent = np.full(N, -1, dtype=np.int64)
    inner = np.nonzero(c0 >= 0)[0]
    parent[c0[inner]] = inner
    parent[c1[inner]] = inner
    codes = np.stack(rows).astype(np.int16)
    ta = TreeArrays(c0, c1, parent, lengths, frees, codes, None, polys)
    return ta, np.asarray(post_index, dtype=np.int64)


class SiteTable:
    """Per-site edit alphabets of one problem: state -> local index 1..m, and
    the normalised priors in CSR form."""

    def __init__(self, Q, normalized=True):
        """Q: per-site dict state -> prior.  normalized: the priors of each
        site sum to 1 over its alphabet (the reference EM model, which
        normalises them, ML_solver.py:69-74); otherwise each site's total mass
        is kept (``qsum``) and used where a missing subtree may be in any edit
        state (the fastEM model, whose alphabets are the observed states only,
        fast_em/phylogeny.py:62-94)."""
        self.K = len(Q)
        self.maps = []
        qoff = [0]
        prior = []
        for k, q in enumerate(Q):
            mass0 = q.get(0, 0.0)
            if mass0 not in (0, 0.0):
                raise ValueError("site %d: the prior puts mass %r on the unedited state 0; "
                                 "the PMM model needs priors over edit states only" % (k, mass0))
            m = {}
            for s in q:
                if s == 0:
                    continue
                m[s] = len(m) + 1
                prior.append(float(q[s]))
            self.maps.append(m)
            qoff.append(len(prior))
        self.qoff = np.asarray(qoff, dtype=np.int32)
        self.prior = np.asarray(prior, dtype=np.float64)
        if normalized:
            self.qsum = np.ones(self.K)
        else:
            self.qsum = np.asarray([self.prior[self.qoff[k]:self.qoff[k + 1]].sum() for k in range(self.K)])

    def encode_row(self, row):
        out = np.empty(len(row), dtype=np.int16)
        for k, x in enumerate(row):
            if x == 0:
                out[k] = 0
            elif x == "?":
                out[k] = CODE_MISSING
            elif x == -1:
                out[k] = CODE_SILENCED
            else:
                try:
                    out[k] = self.maps[k][x]
                except KeyError:
                    raise KeyError("site %d: observed state %r has no prior" % (k, x)) from None
        return out


class ProblemSpec:
    """One problem: trees (as TreeArrays), its site table and phi statistics."""

    def __init__(self, trees, sites, nobs, phimin, flags=LAML_OPT_PHI | LAML_OPT_NU):
        self.trees = trees
        self.sites = sites
        self.nobs = float(nobs)
        self.phimin = float(phimin)
        self.flags = int(flags)


class ForestLayout:
    """Level-ordered host arrays of a forest (no device work)."""

    def __init__(self, problems, dmin=DMIN, dmax=DMAX):
        self.problems = problems
        trees = [t for p in problems for t in p.trees]
        tree_prob = np.concatenate([np.full(len(p.trees), i, dtype=np.int32) for i, p in enumerate(problems)])
        tree_K = np.asarray([p.sites.K for p in problems for _ in p.trees], dtype=np.int32)
        T = len(trees)
        sizes = np.asarray([t.n_nodes for t in trees], dtype=np.int64)
        base = np.concatenate([[0], np.cumsum(sizes)])
        N = int(base[-1])
        height = np.concatenate([t.height for t in trees])
        tree_of_g = np.repeat(np.arange(T, dtype=np.int64), sizes)
        post_idx = np.concatenate([np.arange(s, dtype=np.int64) for s in sizes])
        order = np.lexsort((post_idx, tree_of_g, height))  # level order -> global post id
        inv = np.empty(N, dtype=np.int64)
        inv[order] = np.arange(N, dtype=np.int64)
        self.order = order
        self.inv = inv
        self.tree_base = base

        def gmap(a, t_idx):
            return np.where(a >= 0, a + base[t_idx], -1)

        c0 = np.concatenate([gmap(t.c0, i) for i, t in enumerate(trees)])
        c1 = np.concatenate([gmap(t.c1, i) for i, t in enumerate(trees)])
        par = np.concatenate([gmap(t.parent, i) for i, t in enumerate(trees)])
        free = np.concatenate([t.free for t in trees])
        poly = np.concatenate([t.poly for t in trees])

        def relabel(a):
            a = a[order]
            return np.where(a >= 0, inv[np.maximum(a, 0)], -1).astype(np.int32)

        self.child0 = relabel(c0)
        self.child1 = relabel(c1)
        self.parent = relabel(par)
        self.tree_of = tree_of_g[order].astype(np.int32)
        self.free = free[order].astype(np.uint8)
        self.poly = poly[order]
        hs = height[order]
        H = int(hs.max()) + 1
        self.num_levels = H
        self.level_off = np.searchsorted(hs, np.arange(H + 1)).astype(np.int32)
        self.num_leaves = int(self.level_off[1])
        self.tree_root = np.asarray([inv[base[i] + t.n_nodes - 1] for i, t in enumerate(trees)], dtype=np.int32)
        # leaves (height 0) lead the level order, sorted by tree and then by
        # post-order index, which is the order of each tree's code rows: the
        # leaf codes are the trees' code matrices back to back (no gather;
        # one tree's matrix is shared as a read-only view)
        code_chunks = [np.ascontiguousarray(t.codes, dtype=np.int16).reshape(-1) for t in trees]
        if len(code_chunks) == 1:
            self.codes = code_chunks[0].view()
            self.codes.flags.writeable = False
        else:
            self.codes = np.concatenate(code_chunks) if code_chunks else np.zeros(0, np.int16)
        qoff_parts = []
        prior_parts = []
        qsum_parts = []
        off = 0
        for p in problems:
            for _ in p.trees:
                qoff_parts.append(p.sites.qoff[:-1].astype(np.int64) + off)
                prior_parts.append(p.sites.prior)
                qsum_parts.append(np.asarray(getattr(p.sites, "qsum", np.ones(p.sites.K)), dtype=np.float64))
                off += len(p.sites.prior)
        self.site_qsum = np.ascontiguousarray(np.concatenate(qsum_parts) if qsum_parts else np.zeros(0))
        self.site_qoff = np.concatenate(qoff_parts + [np.asarray([off], dtype=np.int64)]).astype(np.int32)
        self.prior = np.concatenate(prior_parts) if prior_parts else np.zeros(0)
        self.num_nodes = N
        self.num_trees = T
        self.num_problems = len(problems)
        self.tree_prob = tree_prob
        self.tree_K = tree_K
        self.site_base = np.concatenate([[0], np.cumsum(tree_K)]).astype(np.int64)
        self.nobs = np.asarray([p.nobs for p in problems], dtype=np.float64)
        self.phimin = np.asarray([p.phimin for p in problems], dtype=np.float64)
        self.flags = np.asarray([p.flags for p in problems], dtype=np.uint32)
        self.dmin = float(dmin)
        self.dmax = float(dmax)

    def desc(self):
        """C-ABI description pointing into this layout's arrays."""
        def P(a, ct):
            return a.ctypes.data_as(ctypes.POINTER(ct))

        d = _lib.ForestDesc()
        d.num_problems = self.num_problems
        d.num_trees = self.num_trees
        d.num_nodes = self.num_nodes
        d.num_leaves = self.num_leaves
        d.num_levels = self.num_levels
        d.level_off = P(self.level_off, ctypes.c_int32)
        d.child0 = P(self.child0, ctypes.c_int32)
        d.child1 = P(self.child1, ctypes.c_int32)
        d.parent = P(self.parent, ctypes.c_int32)
        d.tree_of = P(self.tree_of, ctypes.c_int32)
        d.free_branch = P(self.free, ctypes.c_uint8)
        d.tree_root = P(self.tree_root, ctypes.c_int32)
        d.tree_problem = P(self.tree_prob, ctypes.c_int32)
        d.tree_nsites = P(self.tree_K, ctypes.c_int32)
        d.site_qoff = P(self.site_qoff, ctypes.c_int32)
        d.prior = P(self.prior, ctypes.c_double)
        d.leaf_codes = P(self.codes, ctypes.c_int16)
        d.problem_nobs = P(self.nobs, ctypes.c_double)
        d.problem_phimin = P(self.phimin, ctypes.c_double)
        d.problem_flags = P(self.flags, ctypes.c_uint32)
        d.dmin = self.dmin
        d.dmax = self.dmax
        d.site_qsum = P(self.site_qsum, ctypes.c_double)
        return d

    def validate(self):
        """Run the C library's host-side validation of the description."""
        lib = _lib.load()
        d = self.desc()
        _lib.check(lib.laml_forest_validate(ctypes.byref(d)))

    # ---------------- parameter layout ----------------
    def lengths_to_level(self, per_tree_lengths):
        """Concatenated post-order branch lengths -> level-order vector."""
        g = np.concatenate(per_tree_lengths)
        return g[self.order]

    def level_to_tree_lengths(self, d_level):
        g = np.empty(self.num_nodes, dtype=np.float64)
        g[self.order] = d_level
        return [g[self.tree_base[i]:self.tree_base[i + 1]] for i in range(self.num_trees)]


class Forest(ForestLayout):
    """Device-resident forest built from ProblemSpecs."""

    def __init__(self, problems, device=0, dmin=DMIN, dmax=DMAX):
        super().__init__(problems, dmin=dmin, dmax=dmax)
        lib = _lib.load()
        self.device = torch.device("cuda", device)
        self.desc_ = self.desc()
        handle = ctypes.c_void_p()
        _lib.check(lib.laml_forest_create(ctypes.byref(self.desc_), int(device), ctypes.byref(handle)))
        self.handle = handle
        self._lib = lib
        N = self.num_nodes
        dev = self.device
        self.counts = torch.zeros(N * 5, dtype=torch.float64, device=dev)
        self.rtilde = torch.zeros(self.num_problems, dtype=torch.float64, device=dev)
        self.nll_buf = torch.zeros(self.num_problems, dtype=torch.float64, device=dev)
        self.site_llh = torch.zeros(int(self.site_base[-1]), dtype=torch.float64, device=dev)
        self.iters_buf = torch.zeros(self.num_problems, dtype=torch.int32, device=dev)

    # ------------------------------------------------------------------
    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._lib.laml_forest_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def set_flags(self, flags):
        a = np.asarray(flags, dtype=np.uint32)
        _lib.check(self._lib.laml_forest_set_flags(self.handle, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))

    # ---------------- ultrametric constraint ----------------
    def set_ultrametric(self, model):
        """Enable the ultrametric M-step with a laml_amd.ultra.HeightModel
        (None disables it)."""
        if model is None:
            _lib.check(self._lib.laml_forest_set_ultrametric(self.handle, None))
            self._ultra = None
            return
        P = self.num_problems
        H = len(model.cls_level_off) - 1
        plo = np.zeros(H * P + 1, dtype=np.int32)
        cp = model.cls_prob
        lvl = np.zeros(model.num_classes, dtype=np.int64)
        for h in range(H):
            lvl[model.cls_level_off[h]:model.cls_level_off[h + 1]] = h
        key = lvl * P + cp
        plo[:] = np.searchsorted(key, np.arange(H * P + 1)).astype(np.int32)
        arrs = dict(plo=plo, up_cls=model.up_cls.astype(np.int32), dn_cls=model.dn_cls.astype(np.int32),
                    up_off=model.up_off.astype(np.float64), dn_off=model.dn_off.astype(np.float64),
                    cls_parent=model.cls_parent.astype(np.int32), cls_pedge=model.cls_pedge.astype(np.int32),
                    inc_off=model.inc_off.astype(np.int32), inc_branch=model.inc_branch.astype(np.int32),
                    inc_sign=model.inc_sign.astype(np.int8), has_free=model.prob_has_free.astype(np.uint8))
        u = _lib.UltraDesc()
        u.num_classes = model.num_classes
        u.num_levels = H
        u.level_problem_off = arrs["plo"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.up_cls = arrs["up_cls"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.dn_cls = arrs["dn_cls"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.up_off = arrs["up_off"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        u.dn_off = arrs["dn_off"].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        u.cls_parent = arrs["cls_parent"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.cls_pedge = arrs["cls_pedge"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.inc_off = arrs["inc_off"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.inc_branch = arrs["inc_branch"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        u.inc_sign = arrs["inc_sign"].ctypes.data_as(ctypes.POINTER(ctypes.c_int8))
        u.problem_has_free = arrs["has_free"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        _lib.check(self._lib.laml_forest_set_ultrametric(self.handle, ctypes.byref(u)))
        self._ultra = (model, arrs)

    def set_heights(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        _lib.check(self._lib.laml_forest_set_heights(self.handle, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def ultra_status(self):
        out = np.zeros(self.num_problems, dtype=np.int32)
        _lib.check(self._lib.laml_forest_ultra_status(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def mstep_status(self):
        """Per-problem status bits of the last M-step (LAML_MSTEP_FAILED_D = 1,
        LAML_MSTEP_FAILED_NU = 2; 0 = solved)."""
        out = np.zeros(self.num_problems, dtype=np.int32)
        _lib.check(self._lib.laml_forest_mstep_status(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    def labels(self):
        """az-partition labels: (leaf rows, internal rows) int16 flat arrays in
        level order, row lengths K of the node's tree."""
        K = self.tree_K[self.tree_of]
        leaf_rows = int(K[: self.num_leaves].sum())
        int_rows = int(K[self.num_leaves:].sum())
        out = np.zeros(leaf_rows + int_rows, dtype=np.int16)
        _lib.check(self._lib.laml_forest_labels(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                                                out.size))
        return out[:leaf_rows], out[leaf_rows:]

    # ---------------- calls (device tensors in, device tensors out) -------
    def nll(self, d, nu, phi, site_llh=False):
        self.state_token = getattr(self, "state_token", 0) + 1
        out = self.nll_buf
        sl = self.site_llh if site_llh else None
        _lib.check(self._lib.laml_nll(self.handle, _lib.ptr(d), _lib.ptr(nu), _lib.ptr(phi), _lib.ptr(out),
                                      _lib.ptr(sl), self.stream()))
        return out

    def estep(self, d, nu, phi):
        _lib.check(self._lib.laml_estep(self.handle, _lib.ptr(d), _lib.ptr(nu), _lib.ptr(phi),
                                        _lib.ptr(self.counts), _lib.ptr(self.rtilde), _lib.ptr(self.nll_buf),
                                        self.stream()))
        return self.counts, self.rtilde, self.nll_buf

    def outside(self, phi):
        _lib.check(self._lib.laml_outside(self.handle, _lib.ptr(phi), _lib.ptr(self.counts), _lib.ptr(self.rtilde),
                                          self.stream()))
        return self.counts, self.rtilde

    def grad(self, d, nu, phi):
        """Gradient of each problem's NLL w.r.t. d (level order), nu and phi,
        by the outside pass over the inside state of the preceding nll()."""
        gd = torch.empty_like(d)
        gnu = torch.empty_like(nu)
        gphi = torch.empty_like(phi)
        _lib.check(self._lib.laml_grad(self.handle, _lib.ptr(d), _lib.ptr(nu), _lib.ptr(phi), _lib.ptr(gd),
                                       _lib.ptr(gnu), _lib.ptr(gphi), self.stream()))
        return gd, gnu, gphi

    def set_checkpoint(self, levels):
        """Keep inside vectors only at level-block boundaries (0: keep all)."""
        _lib.check(self._lib.laml_forest_set_checkpoint(self.handle, int(levels)))
        self.state_token = getattr(self, "state_token", 0) + 1

    def inside_rows(self):
        return int(self._lib.laml_forest_inside_rows(self.handle))

    def posteriors(self, phi):
        """(post0, postS): P(node unedited | D) and P(node silenced | D) per
        (node, site) on the state of the preceding estep, as (leaf rows,
        internal rows) flat arrays in level order like ``labels``."""
        K = self.tree_K[self.tree_of]
        leaf_rows = int(K[: self.num_leaves].sum())
        int_rows = int(K[self.num_leaves:].sum())
        p0 = np.zeros(leaf_rows + int_rows)
        pS = np.zeros(leaf_rows + int_rows)
        f64 = ctypes.POINTER(ctypes.c_double)
        _lib.check(self._lib.laml_posteriors(self.handle, _lib.ptr(phi), p0.ctypes.data_as(f64),
                                             pS.ctypes.data_as(f64), p0.size, self.stream()))
        return (p0[:leaf_rows], p0[leaf_rows:]), (pS[:leaf_rows], pS[leaf_rows:])

    def quartet_em(self, qnodes, len_in, opt, llh_const, nu, phi, max_iter, tol):
        """Quartet-local EM of NNI candidates on this forest's resident
        inside/outside state (see laml_quartet_em).  qnodes [n, 5] (p, s, A,
        W, C) level-order node ids; len_in [n, 5]; opt [n] bit masks;
        llh_const [n].  Returns (lengths [n, 5], nll [n], iters [n]) device
        tensors."""
        dev = self.device
        q = torch.as_tensor(np.ascontiguousarray(qnodes, dtype=np.int32), device=dev)
        li = torch.as_tensor(np.ascontiguousarray(len_in, dtype=np.float64), device=dev)
        om = torch.as_tensor(np.ascontiguousarray(opt, dtype=np.uint8), device=dev)
        lc = torch.as_tensor(np.ascontiguousarray(llh_const, dtype=np.float64), device=dev)
        n = int(q.shape[0])
        lo = torch.empty((n, 5), dtype=torch.float64, device=dev)
        nl = torch.empty(n, dtype=torch.float64, device=dev)
        it = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.check(self._lib.laml_quartet_em(self.handle, n, _lib.ptr(q), _lib.ptr(li), _lib.ptr(om), _lib.ptr(lc),
                                             _lib.ptr(nu), _lib.ptr(phi), int(max_iter), float(tol), _lib.ptr(lo),
                                             _lib.ptr(nl), _lib.ptr(it), self.stream()))
        return lo, nl, it

    def set_pass_mode(self, mode):
        """0 (levels), 1 (tree walk: one launch per pass, a block per (tree,
        site chunk)) or 2 (hybrid: level launches below the top level of the
        launch shapes, one tree-walk launch above)."""
        _lib.check(self._lib.laml_forest_set_pass_mode(self.handle, int(mode)))

    def pass_mode(self):
        return int(self._lib.laml_forest_pass_mode(self.handle))

    def set_fusion(self, on):
        """Cherry fusion of the level passes (off by default): the nodes
        whose children are both leaves are formed and consumed by their
        parents.  posteriors() and quartet_em() need an E-step run with it
        off."""
        _lib.check(self._lib.laml_forest_set_fusion(self.handle, 1 if on else 0))

    def fusion(self):
        """Whether the next level pass fuses the cherries."""
        return bool(self._lib.laml_forest_fusion(self.handle))

    def set_chains(self, which, inside_top=None, outside_top=None):
        """Heavy-path chains for the levels below a chain-region top (and the
        split): one wavefront walks a node and its successive heavy children
        at a site chunk, keeping their states in registers (same bits as
        level launches with one chunk per wavefront).  which: True (both
        passes), False, or the flags CHAINS_OUTSIDE | CHAINS_INSIDE;
        inside_top / outside_top: region top level per pass (0: up to the
        split; None keeps the current value)."""
        flags = (_lib.CHAINS_OUTSIDE | _lib.CHAINS_INSIDE) if which is True else int(which)
        _, ti, to = self.get_chains()
        ti = ti if inside_top is None else int(inside_top)
        to = to if outside_top is None else int(outside_top)
        _lib.check(self._lib.laml_forest_set_chains(self.handle, flags, ti, to))

    def get_chains(self):
        """(requested flags, inside region top, outside region top)."""
        w, ti, to = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self._lib.laml_forest_get_chains(self.handle, ctypes.byref(w), ctypes.byref(ti), ctypes.byref(to)))
        return w.value, ti.value, to.value

    def chains(self):
        """The chain flags the next passes use (0: none)."""
        return int(self._lib.laml_forest_chains(self.handle))

    def get_shapes(self):
        """Launch shapes: (inside_cpw, outside_cpw, top_level); the cpw arrays
        hold, per level, the 64-site chunks one wavefront walks (all 0 when
        the rows hold at most 32 sites)."""
        H = self.num_levels
        a = np.zeros(H, dtype=np.int32)
        b = np.zeros(H, dtype=np.int32)
        top = np.zeros(1, dtype=np.int32)
        i32 = ctypes.POINTER(ctypes.c_int32)
        _lib.check(self._lib.laml_forest_get_shapes(self.handle, a.ctypes.data_as(i32), b.ctypes.data_as(i32),
                                                    top.ctypes.data_as(i32)))
        return a, b, int(top[0])

    def set_shapes(self, inside_cpw=None, outside_cpw=None, top_level=None):
        """Set launch shapes (None keeps the current value)."""
        a0, b0, t0 = self.get_shapes()
        a = np.ascontiguousarray(a0 if inside_cpw is None else inside_cpw, dtype=np.int32)
        b = np.ascontiguousarray(b0 if outside_cpw is None else outside_cpw, dtype=np.int32)
        if a.shape != (self.num_levels,) or b.shape != (self.num_levels,):
            raise ValueError("shape arrays need one entry per level")
        i32 = ctypes.POINTER(ctypes.c_int32)
        _lib.check(self._lib.laml_forest_set_shapes(self.handle, a.ctypes.data_as(i32), b.ctypes.data_as(i32),
                                                    int(t0 if top_level is None else top_level)))

    def node_problem(self):
        """Problem index of every node (device int64 tensor, cached)."""
        if getattr(self, "_node_prob_t", None) is None:
            self._node_prob_t = torch.as_tensor(self.tree_prob[self.tree_of].astype(np.int64), device=self.device)
        return self._node_prob_t

    def mstep(self, d, nu, phi, counts=None, rtilde=None):
        c = self.counts if counts is None else counts
        r = self.rtilde if rtilde is None else rtilde
        _lib.check(self._lib.laml_mstep(self.handle, _lib.ptr(c), _lib.ptr(r), _lib.ptr(d), _lib.ptr(nu),
                                        _lib.ptr(phi), self.stream()))

    def em(self, d, nu, phi, max_iter, tol, trace=None):
        _lib.check(self._lib.laml_em(self.handle, _lib.ptr(d), _lib.ptr(nu), _lib.ptr(phi), int(max_iter),
                                     float(tol), _lib.ptr(self.nll_buf), _lib.ptr(self.iters_buf),
                                     _lib.ptr(trace), self.stream()))
        return self.nll_buf, self.iters_buf
