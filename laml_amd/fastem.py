"""fastEM surface: the reference's fastEM encodings and its EM optimiser API,
with the EM itself run by the exact FP64 device engine.

``parse_tree`` and ``parse_data`` restate laml_libs/fastEM_solver.py:17-146 and
fast_em/phylogeny.py:48-114 (integer work, bit-exact): nodes relabelled
leaves first (in ``ordered_leaf_labels`` order, else preorder) then internal
nodes in preorder, a synthetic out-degree-1 root dropped, each site's
observed non-zero states recoded 1..k in sorted order ('?' -> -1), the priors
restricted to the observed states (not renormalised).

``EMOptimizer`` keeps the call of fast_em/laml.py:196-276
(``optimize(tree, branch_lengths, nu, phi, branch_mask, fixed_nu, fixed_phi)``
-> dict with nllh, branch_lengths, nu, phi, em_iterations) on the fastEM
model: each site's alphabet is its observed states with their raw prior mass,
and -1 entries are missing like '?'.  The reference solves that model in FP32
with a joint L-BFGS surrogate M-step (fast_em/laml.py:62-110) and EPS clamps
(fast_em/calculations.py:9); here the E-step is exact FP64 and the M-step is
the exact EM M-step, stopped by the same relative criterion (1e-5,
fast_em/laml.py:46,183).  The returned parameters are those of the returned
likelihood (the reference returns the previous iterate's with the last
likelihood).

``EMOptimizer(..., m_step="surrogate")`` runs the reference's own EM instead
(fast_em/laml.py:112-194) in FP64 without the EPS clamps: device E-steps, and
an M-step that minimises the reference's joint surrogate (M_step_loss_fn,
fast_em/laml.py:62-78: the expected complete-data log-likelihood over log
branch lengths and logit nu, phi, with the raw expected counts) by 100
L-BFGS iterations on the host, its stopping rule, and its return values
(the parameters before the last M-step, with the likelihood after it).
"""
import random as _random

import numpy as np
import torch

from ._lib import CODE_MISSING, LAML_OPT_NU, LAML_OPT_PHI
from .forest import Forest, ProblemSpec, SiteTable, TreeArrays

# fast_em/calculations.py:9, fast_em/laml.py:46
EPS = 1e-6
EM_STOPPING_CRITERION = 1e-5
# branch-length box of the M-step on the fastEM model (its branch lengths are
# exp of an unconstrained parameter, floored at EPS)
FASTEM_DMIN = EPS
FASTEM_DMAX = 1e3
MAX_EM_ITER = 10000
# L-BFGS iterations of one surrogate M-step (fast_em/laml.py:45)
M_STEP_DESCENT_STEPS = 100

try:  # the reference hands networkx graphs around; plain arrays work without it
    import networkx as _nx
except ImportError:  # pragma: no cover
    _nx = None


class ParsedTree(dict):
    """parse_tree's result: 'parent' (int array by new index, -1 at the root),
    'children', 'num_leaves', 'branch_lengths' (float32), 'branch_mask'
    (bool, True = optimise), 'relabeling_vector' (labels by new index) and,
    when networkx is importable, 'nxtree' (DiGraph on the new indices)."""


def parse_tree(tree, has_branch_mask=False, ordered_leaf_labels=None):
    """fastEM_solver.parse_tree (fastEM_solver.py:46-146).  Unlabelled nodes
    get the label ``node_<preorder index>`` (set on the tree, as there); a
    missing branch length is drawn from random.uniform(0.01, 0.5)."""
    pre = list(tree.traverse_preorder())
    labels = []
    for idx, v in enumerate(pre):
        if v.label is None:
            v.label = "node_%d" % idx
        labels.append(v.label)
    index = {id(v): i for i, v in enumerate(pre)}
    parent = [index[id(v.parent)] if v.parent is not None else -1 for v in pre]
    # a node's children in the order the preorder reaches them (last child
    # first): the order of its successors in the reference's DiGraph, whose
    # edges are added as the preorder reaches each child
    children = [[index[id(c)] for c in reversed(v.children)] for v in pre]
    keep = list(range(len(pre)))
    if len(children[0]) == 1:  # synthetic root with out-degree 1
        keep = keep[1:]
        parent[children[0][0]] = -1
    leaves = [i for i in keep if not children[i]]
    if ordered_leaf_labels is not None:
        by_label = {labels[i]: i for i in leaves}
        leaf_order = [by_label[x] for x in ordered_leaf_labels if x in by_label]
    else:
        leaf_order = leaves
    internal = [i for i in keep if children[i]]
    order = leaf_order + internal
    new = {old: k for k, old in enumerate(order)}
    bl, mask = [], []
    for old in order:
        v = pre[old]
        b = v.edge_length if v.edge_length is not None else _random.uniform(0.01, 0.5)
        bl.append(b)
        mask.append((not getattr(v, "mark_fixed", False)) if has_branch_mask else True)
    out = ParsedTree()
    out["parent"] = np.asarray([new[parent[o]] if parent[o] >= 0 else -1 for o in order], dtype=np.int64)
    out["children"] = [[new[c] for c in children[o]] for o in order]
    out["num_leaves"] = len(leaf_order)
    out["branch_lengths"] = np.asarray(bl, dtype=np.float32)
    out["branch_mask"] = np.asarray(mask, dtype=bool)
    out["relabeling_vector"] = [labels[o] for o in order]
    if _nx is not None:
        g = _nx.DiGraph()
        for k, o in enumerate(order):
            attrs = dict(label=labels[o], branch_length=pre[o].edge_length)
            if has_branch_mask:
                attrs["branch_mask"] = getattr(pre[o], "mark_fixed", False)
            g.add_node(k, **attrs)
        for k, o in enumerate(order):
            for c in out["children"][k]:
                g.add_edge(k, c)
        out["nxtree"] = g
    return out


class Phylogeny:
    """build_phylogeny's result (fast_em/phylogeny.py:48-114)."""

    def __init__(self, num_leaves, num_characters, max_alphabet_size, root, parsed, character_matrix,
                 mutation_priors, state_mappings):
        self.num_leaves = num_leaves
        self.num_characters = num_characters
        self.max_alphabet_size = max_alphabet_size
        self.root = root
        self.tree = parsed.get("nxtree", parsed)
        self.character_matrix = character_matrix    # int32 [num_leaves, K], row = leaf index
        self.mutation_priors = mutation_priors      # float32 [K, max_alphabet_size]
        self.state_mappings = state_mappings        # per site: new index -> original state


def parse_data(parsed, data, prior):
    """fastEM_solver.parse_data + build_phylogeny: recode the character matrix
    in the parsed tree's leaf order and restrict the priors to the observed
    states."""
    n = parsed["num_leaves"]
    leaf_labels = parsed["relabeling_vector"][:n]
    charMtrx = data["charMtrx"]
    Q = prior["Q"]
    K = len(next(iter(charMtrx.values())))
    rows = np.asarray([[(-1 if x == "?" else int(x)) for x in charMtrx[c]] for c in leaf_labels], dtype=np.int64)
    rec = np.empty_like(rows)
    mappings = []
    inv = []
    max_k = 0
    for k in range(K):
        col = rows[:, k]
        valid = sorted(set(int(s) for s in np.unique(col)) - {-1, 0})
        mp = {s: j + 1 for j, s in enumerate(valid)}
        mappings.append(mp)
        inv.append({j + 1: s for j, s in enumerate(valid)})
        max_k = max(max_k, len(valid))
        rec[:, k] = [mp.get(int(x), int(x)) for x in col]
    pri = np.zeros((K, max_k), dtype=np.float32)
    for k in range(K):
        for s, j in mappings[k].items():
            if s not in Q[k]:
                raise KeyError("missing mutation prior for character %d, state %r" % (k, s))
            pri[k, j - 1] = Q[k][s]
    root = int(np.nonzero(parsed["parent"] < 0)[0][0])
    return Phylogeny(n, K, max_k, root, parsed, rec.astype(np.int32), pri, inv)


def _postorder(parent, children, root):
    out = []
    stack = [(root, False)]
    while stack:
        v, done = stack.pop()
        if done:
            out.append(v)
            continue
        stack.append((v, True))
        for c in reversed(children[v]):
            stack.append((c, False))
    return out


def fastem_problem(children, root, branch_lengths, branch_mask, character_matrix, mutation_priors, flags):
    """ProblemSpec of the fastEM model for one tree given by children lists
    on node indices (leaves 0..n-1 = character-matrix rows)."""
    post = _postorder(None, children, root)
    pos = {v: i for i, v in enumerate(post)}
    N = len(post)
    c0 = np.full(N, -1, dtype=np.int64)
    c1 = np.full(N, -1, dtype=np.int64)
    par = np.full(N, -1, dtype=np.int64)
    length = np.zeros(N)
    free = np.zeros(N, dtype=np.uint8)
    leaf_rows = []
    for i, v in enumerate(post):
        ch = children[v]
        if ch:
            if len(ch) != 2:
                raise ValueError("the fastEM model needs a binary tree")
            c0[i], c1[i] = pos[ch[0]], pos[ch[1]]
            par[pos[ch[0]]] = i
            par[pos[ch[1]]] = i
        else:
            leaf_rows.append(v)
        length[i] = float(branch_lengths[v])
        free[i] = 1 if bool(branch_mask[v]) else 0
    C = np.asarray(character_matrix)[leaf_rows]
    codes = np.where(C < 0, CODE_MISSING, C).astype(np.int16)
    K, A = np.asarray(mutation_priors).shape
    Qd = [{j + 1: float(mutation_priors[k][j]) for j in range(A)} for k in range(K)]
    sites = SiteTable(Qd, normalized=False)
    ta = TreeArrays(c0, c1, par, length, free, codes)
    nobs = float((codes != CODE_MISSING).sum())
    phimin = 1.0 / (K * codes.shape[0])
    return ProblemSpec([ta], sites, nobs, phimin, flags), post


def _log1mexp(x):
    """log(1 - e^-x) for x > 0 (fast_em/laml.py:54-57)."""
    return np.where(x > np.log(2.0), np.log1p(-np.exp(-np.minimum(x, 1e300))), np.log(-np.expm1(-x)))


def surrogate_loss(x, S, n_missing, n_not_missing, rt):
    """The reference's M-step loss (M_step_loss_fn, fast_em/laml.py:62-78)
    and its gradient, in FP64 without clamps.  x = (log branch lengths [N],
    logit nu, logit phi); S: raw expected counts S0..S4 per branch [N, 5]
    (summed over sites); rt: sum over missing leaf entries of P(not silenced
    | D), i.e. num_missing minus the summed leaf responsibilities."""
    N = S.shape[0]
    d = np.exp(x[:N])
    nu = 1.0 / (1.0 + np.exp(-x[N]))
    phi = 1.0 / (1.0 + np.exp(-x[N + 1]))
    dn = d * nu
    l1 = _log1mexp(d)
    l2 = _log1mexp(dn)
    B = S[:, 2] + S[:, 4]
    q = (-S[:, 0] * d * (1.0 + nu) + S[:, 1] * (l1 - d * nu) + B * l2 - S[:, 3] * d * nu)
    c6 = n_not_missing * np.log1p(-phi) + rt * np.log(phi)
    loss = -(q.sum() + c6)
    # derivatives of -loss
    gd = -S[:, 0] * (1.0 + nu) + S[:, 1] * (1.0 / np.expm1(d) - nu) + B * nu / np.expm1(dn) - S[:, 3] * nu
    gnu = np.sum(d * (-S[:, 0] - S[:, 1] - S[:, 3] + B / np.expm1(dn)))
    gphi = -n_not_missing / (1.0 - phi) + rt / phi
    g = np.empty(N + 2)
    g[:N] = -gd * d
    g[N] = -gnu * nu * (1.0 - nu)
    g[N + 1] = -gphi * phi * (1.0 - phi)
    return loss, g


def surrogate_mstep(x0, free, S, n_missing, n_not_missing, rt, steps=M_STEP_DESCENT_STEPS):
    """One surrogate M-step (fast_em/laml.py:101-110): `steps` L-BFGS
    iterations over the free entries of x (the reference masks the gradient
    of the fixed ones, so they keep their values)."""
    from scipy.optimize import minimize

    idx = np.nonzero(free)[0]

    def f(z):
        x = x0.copy()
        x[idx] = z
        loss, g = surrogate_loss(x, S, n_missing, n_not_missing, rt)
        return loss, g[idx]

    res = minimize(f, x0[idx], jac=True, method="L-BFGS-B",
                   options={"maxiter": steps, "maxcor": 10, "ftol": 0.0, "gtol": 0.0})
    x = x0.copy()
    x[idx] = res.x
    return x


class EMOptimizer:
    """fast_em.laml.EMOptimizer (fast_em/laml.py:196-276) on the device.
    m_step: "exact" (the default) solves each M-step exactly with the
    device's EM M-step; "surrogate" runs the reference's joint L-BFGS
    surrogate M-step (see surrogate_mstep)."""

    def __init__(self, mutation_priors, character_matrix, verbose=False, m_step="exact"):
        if m_step not in ("exact", "surrogate"):
            raise ValueError("m_step must be 'exact' or 'surrogate'")
        self.mutation_priors = np.asarray(mutation_priors, dtype=np.float64)
        self.character_matrix = np.asarray(character_matrix)
        self.num_leaves = self.character_matrix.shape[0]
        self.alphabet_size = self.mutation_priors.shape[1]
        self.num_characters = self.character_matrix.shape[1]
        self.verbose = verbose
        self.m_step = m_step

    def optimize(self, tree, branch_lengths, nu, phi, branch_mask, fixed_nu=False, fixed_phi=False):
        if isinstance(tree, dict) and "children" in tree:
            children = tree["children"]
        else:  # networkx DiGraph on node indices
            children = [list(tree.successors(i)) for i in range(tree.number_of_nodes())]
        has_parent = np.zeros(len(children), dtype=bool)
        for ch in children:
            for c in ch:
                has_parent[c] = True
        root = int(np.nonzero(~has_parent)[0][0])
        flags = (0 if fixed_nu else LAML_OPT_NU) | (0 if fixed_phi else LAML_OPT_PHI)
        bl = np.maximum(np.asarray(branch_lengths, dtype=np.float64), EPS)
        spec, post = fastem_problem(children, root, bl, np.asarray(branch_mask), self.character_matrix,
                                    self.mutation_priors, flags)
        if not torch.cuda.is_available():
            raise RuntimeError("EMOptimizer needs a ROCm GPU")
        F = Forest([spec], device=torch.cuda.current_device(), dmin=FASTEM_DMIN, dmax=FASTEM_DMAX)
        if self.m_step == "surrogate":
            try:
                return self._optimize_surrogate(F, spec, post, np.asarray(branch_lengths, dtype=np.float64),
                                                np.asarray(branch_mask, dtype=bool), nu, phi, fixed_nu, fixed_phi,
                                                len(children))
            finally:
                F.close()
        try:
            d = torch.tensor(F.lengths_to_level([spec.trees[0].length]), dtype=torch.float64, device=F.device)
            clamp = lambda x: min(max(float(x), EPS), 1.0 - EPS)
            nu_t = torch.tensor([clamp(nu)], dtype=torch.float64, device=F.device)
            phi_t = torch.tensor([clamp(phi)], dtype=torch.float64, device=F.device)
            nll, iters = F.em(d, nu_t, phi_t, MAX_EM_ITER, EM_STOPPING_CRITERION)
            lengths_post = F.level_to_tree_lengths(d.cpu().numpy())[0]
            out_bl = np.empty(len(children))
            for i, v in enumerate(post):
                out_bl[v] = lengths_post[i]
            res = {"nllh": float(nll[0].item()), "branch_lengths": out_bl, "nu": float(nu_t.item()),
                   "phi": float(phi_t.item()), "em_iterations": int(iters[0].item())}
        finally:
            F.close()
        return res

    def _optimize_surrogate(self, F, spec, post, branch_lengths, branch_mask, nu, phi, fixed_nu, fixed_phi, n):
        """optimize_parameters_expectation_maximization (fast_em/laml.py:112-194)
        in FP64: device E-steps and likelihoods, host surrogate M-steps."""
        N = spec.trees[0].n_nodes
        dev = F.device
        # the start: lengths as given (positive), nu and phi inside (0, 1)
        # (fast_em/laml.py:128-132 clamp them at EPS; FP64 needs no clamp
        # beyond keeping the logit finite)
        lp = np.asarray(spec.trees[0].length, dtype=np.float64)  # post-order of the device tree
        if np.any(lp <= 0.0):
            raise ValueError("surrogate M-step needs positive branch lengths")
        m = np.asarray([min(max(float(nu), 1e-300), 1.0 - 1e-16), min(max(float(phi), 1e-300), 1.0 - 1e-16)])
        x = np.concatenate([np.log(lp), np.log(m / (1.0 - m))])
        free = np.concatenate([spec.trees[0].free.astype(bool), [not fixed_nu, not fixed_phi]])
        C = self.character_matrix
        n_missing = float((C == -1).sum())
        n_not_missing = float((C != -1).sum())

        def to_dev(xv):
            d = torch.tensor(F.lengths_to_level([np.exp(xv[:N])]), dtype=torch.float64, device=dev)
            mp = 1.0 / (1.0 + np.exp(-xv[N:]))
            return (d, torch.tensor([mp[0]], dtype=torch.float64, device=dev),
                    torch.tensor([mp[1]], dtype=torch.float64, device=dev))

        def nllh(xv):
            return float(F.nll(*to_dev(xv)).item())

        previous = nllh(x)
        iteration = 0
        while True:
            iteration += 1
            counts, rt, _ = F.estep(*to_dev(x))
            S = np.empty((N, 5))
            S[F.order] = counts.view(-1, 5).cpu().numpy()
            x_new = surrogate_mstep(x, free, S, n_missing, n_not_missing, float(rt.item()))
            current = nllh(x_new)
            if abs(current - previous) / abs(previous) < EM_STOPPING_CRITERION:
                break
            if np.isnan(current):
                current = previous
                break
            x = x_new
            previous = current
        # the reference returns the parameters before the last M-step with the
        # likelihood after it (fast_em/laml.py:181-194)
        out_bl = np.empty(n)
        for i, v in enumerate(post):
            out_bl[v] = np.exp(x[i])
        mp = 1.0 / (1.0 + np.exp(-x[N:]))
        self.last_x = x
        self.last_x_new = x_new
        return {"nllh": current, "branch_lengths": out_bl, "nu": float(mp[0]), "phi": float(mp[1]),
                "em_iterations": iteration}
