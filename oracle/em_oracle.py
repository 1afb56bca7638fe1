"""CPU oracle for LAML's PMM likelihood and EM path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product (``laml_amd``) never imports, links or runs it.

This is a scalar, FP64, log-space restatement of the reference's ``EM_solver``
arithmetic (raphael-group/LAML v1.0.5), following its steps function by
function so that the GPU path (which uses a different, linear-space compressed
recursion) is checked against an independent formulation:

  az_partition            laml_libs/ML_solver.py:125-154
  estep_in  (L0/L1)       laml_libs/EM_solver.py:91-133
  lineage_llh             laml_libs/EM_solver.py:135-141
  negative_llh            laml_libs/ML_solver.py:252-254
  estep_out (out0/out1)   laml_libs/EM_solver.py:143-204
  _out_alpha_up           laml_libs/EM_solver.py:206-245
  estep_posterior (S0-4)  laml_libs/EM_solver.py:247-300
  mstep                   laml_libs/EM_solver.py:307-459
  em_optimization         laml_libs/EM_solver.py:461-512
  optimize_one            laml_libs/EM_solver.py:514-527
  optimize (initials)     laml_libs/ML_solver.py:256-336
  ini_all / x2params      laml_libs/ML_solver.py:195-247, EM_solver.py:45-51
  polytomy marking        laml_libs/EM_solver.py:26-43

The reference's M-step calls MOSEK through cvxpy (EM_solver.py:349-407), a
closed third-party solver.  The oracle replaces it with exact solutions of the
same concave problems: per-branch bisection on the stationarity condition
(the problem is separable without the ultrametric constraint) and bisection
for the 1-D nu problem.  Those solutions are pinned by the KKT conditions in
tests, not by the reference.  With the ultrametric constraint (the reference's
equality system, EM_solver.py:53-89, restated in ``ultrametric_constr``) the
branch problem is solved by scipy's SLSQP in branch-length space; its accuracy
(~1e-8) is what the device solver is compared against.

Tree handling restates what the reference takes from treeswift (>=1.1.39, not
installed here): Newick reading, post/pre-order traversal with children in
Newick order, ``suppress_unifurcations`` and ``resolve_polytomies``.
"""
import math
import random as _random

# laml_libs/core_constants.py:3-8
MIN_LLH = -float("inf")
EPS = 1e-10
CONV_EPS = 1e-8
DMIN = 0.005
DMAX = 10.0


# --------------------------------------------------------------------------
# minimal tree (treeswift semantics)
# --------------------------------------------------------------------------
class ONode:
    __slots__ = ("label", "edge_length", "children", "parent", "a")

    def __init__(self, label=None, edge_length=None):
        self.label = label
        self.edge_length = edge_length
        self.children = []
        self.parent = None
        self.a = {}

    def is_leaf(self):
        return len(self.children) == 0

    def is_root(self):
        return self.parent is None

    def add_child(self, c):
        self.children.append(c)
        c.parent = self


def parse_newick(s):
    """Parse a Newick string into an ONode tree (labels as str, lengths float)."""
    s = s.strip()
    if s.endswith(";"):
        s = s[:-1]
    root = ONode()
    cur = root
    i = 0
    n = len(s)
    stack = []
    while i < n:
        ch = s[i]
        if ch == "(":
            nn = ONode()
            cur.add_child(nn)
            stack.append(cur)
            cur = nn
            i += 1
        elif ch == ",":
            par = stack[-1]
            nn = ONode()
            par.add_child(nn)
            cur = nn
            i += 1
        elif ch == ")":
            cur = stack.pop()
            i += 1
        elif ch == ":":
            j = i + 1
            while j < n and s[j] not in ",():;":
                j += 1
            cur.edge_length = float(s[i + 1:j])
            i = j
        else:
            j = i
            while j < n and s[j] not in ",():;":
                j += 1
            cur.label = s[i:j]
            i = j
    return root


def postorder(root):
    out = []
    stack = [(root, False)]
    while stack:
        v, done = stack.pop()
        if done:
            out.append(v)
        else:
            stack.append((v, True))
            for c in reversed(v.children):
                stack.append((c, False))
    return out


def preorder(root):
    out = []
    stack = [root]
    while stack:
        v = stack.pop()
        out.append(v)
        for c in reversed(v.children):
            stack.append(c)
    return out


def suppress_unifurcations(root):
    """treeswift Tree.suppress_unifurcations: splice out 1-child nodes."""
    q = [root]
    while q:
        node = q.pop(0)
        if len(node.children) != 1:
            q.extend(node.children)
            continue
        child = node.children.pop()
        if node.parent is None:
            root = child
            child.parent = None
        else:
            par = node.parent
            par.children.remove(node)
            par.add_child(child)
        if node.edge_length is not None:
            if child.edge_length is None:
                child.edge_length = 0.0
            child.edge_length += node.edge_length
        if child.label is None and node.label is not None:
            child.label = node.label
        q.append(child)
    return root


def resolve_polytomies(root):
    """treeswift Tree.resolve_polytomies: pop the last two children into a new
    zero-length child until every node has at most two children."""
    new_nodes = []
    q = [root]
    while q:
        node = q.pop(0)
        while len(node.children) > 2:
            c1 = node.children.pop()
            c2 = node.children.pop()
            nn = ONode(edge_length=0.0)
            node.add_child(nn)
            nn.add_child(c1)
            nn.add_child(c2)
            new_nodes.append(nn)
        q.extend(node.children)
    return new_nodes


def to_newick(root):
    """Newick text of an ONode tree (lengths written with repr)."""
    def rec(v):
        out = ""
        if v.children:
            out = "(" + ",".join(rec(c) for c in v.children) + ")"
        if v.label is not None:
            out += str(v.label)
        if v.edge_length is not None:
            out += ":" + repr(float(v.edge_length))
        return out
    return rec(root) + ";"


def find_lca(root, a, b):
    """Lowest common ancestor of the nodes labelled a and b
    (laml_libs/lca_lib.py:13-95 answers the same query with an Euler tour
    and a segment tree; here by climbing from the deeper node)."""
    by_label = {}
    depth = {}
    for v in preorder(root):
        depth[v] = 0 if v.parent is None else depth[v.parent] + 1
        if v.label is not None:
            by_label[v.label] = v
    x, y = by_label[a], by_label[b]
    while depth[x] > depth[y]:
        x = x.parent
    while depth[y] > depth[x]:
        y = y.parent
    while x is not y:
        x, y = x.parent, y.parent
    return x


def log_sum_exp(xs):
    """laml_libs/EM_solver.py:8-13 (all -inf inputs give NaN, as there)."""
    m = max(xs)
    return m + math.log(sum(math.exp(x - m) for x in xs))


def pseudo_log(x):
    """laml_libs/EM_solver.py:15-16."""
    return math.log(x) if x > 0 else MIN_LLH


# --------------------------------------------------------------------------
# the EM oracle
# --------------------------------------------------------------------------
class OracleEM:
    """Restatement of EM_solver for one or more trees over one character matrix."""

    def __init__(self, tree_strs, charMtrx, Q, nu=0.0, phi=0.0):
        # ML_solver.__init__ (ML_solver.py:19-86)
        self.trees = []
        for t in tree_strs:
            r = parse_newick(t) if isinstance(t, str) else t
            r = suppress_unifurcations(r)
            self.trees.append(r)
        cols = list(zip(*charMtrx.values()))
        uninf = [i for i, c in enumerate(cols) if all(v in (0, -1, "?") for v in c)]
        if len(Q) != len(cols):
            if len(uninf) == max(len(Q), len(cols)) - min(len(Q), len(cols)):
                keep = [i for i in range(len(cols)) if i not in uninf]
                charMtrx = {c: [v[i] for i in keep] for c, v in charMtrx.items()}
        self.charMtrx = charMtrx
        self.Q = []
        for q in Q:
            s = sum(q[x] for x in q)
            self.Q.append({x: q[x] / s for x in q})
        self.nu = nu
        self.phi = phi
        self.numsites = len(self.charMtrx[next(iter(self.charMtrx))])
        self.dmin = DMIN
        self.dmax = DMAX
        # EM_solver.__init__ / __mark_polytomies__ (EM_solver.py:19-43)
        self.num_polytomy_mark = 0
        for r in self.trees:
            for v in preorder(r):
                v.a["poly"] = False
            new = resolve_polytomies(r)
            for nn in new:
                nn.a["poly"] = True
                nn.edge_length = self.dmin * 0.01
                self.num_polytomy_mark += 1
        self.num_edges = sum(len(postorder(r)) for r in self.trees)
        for r in self.trees:
            for v in postorder(r):
                v.a["fixed"] = False

    # ---------------- parameters ----------------
    def ini_all(self, rng, fixed_phi=None, fixed_nu=None):
        """ML_solver.ini_brlens/ini_nu/ini_phi/ini_all (ML_solver.py:195-213)."""
        x = [rng.random() * (self.dmax / 2 - 2 * self.dmin) + 2 * self.dmin
             for _ in range(self.num_edges)]
        idx = 0
        for r in self.trees:
            for v in postorder(r):
                if v.edge_length is not None:
                    x[idx] = max(2 * self.dmin, v.edge_length)
                idx += 1
        nu = rng.random() * 0.99 if fixed_nu is None else fixed_nu
        phi = rng.random() * 0.99 if fixed_phi is None else fixed_phi
        return x + [nu, phi]

    def x2params(self, x, fixed_nu=None, fixed_phi=None):
        """EM_solver.x2brlen (EM_solver.py:45-51) + ML_solver.x2nu/x2phi."""
        i = 0
        for r in self.trees:
            for v in postorder(r):
                if not v.a["poly"]:
                    v.edge_length = x[i]
                    i += 1
        self.nu = x[self.num_edges] if fixed_nu is None else fixed_nu
        self.phi = x[self.num_edges + 1] if fixed_phi is None else fixed_phi

    def branch_lengths(self):
        return [v.edge_length for r in self.trees for v in postorder(r)]

    # ---------------- az partition ----------------
    def az_partition(self):
        """ML_solver.az_partition (ML_solver.py:125-154)."""
        K = self.numsites
        for r in self.trees:
            for v in postorder(r):
                if v.is_leaf():
                    row = self.charMtrx[v.label]
                    al = []
                    for k in range(K):
                        c = row[k]
                        if c == 0:
                            al.append("z")
                        elif c == -1:
                            al.append("?")
                        else:
                            al.append(c)
                    v.a["alpha"] = al
                else:
                    al = []
                    for k in range(K):
                        S = set(c.a["alpha"][k] for c in v.children)
                        R = S - {"z", "?", -1}
                        if "z" in S or len(R) > 1:
                            al.append("z")
                        elif len(R) == 1:
                            al.append(next(iter(R)))
                        else:
                            al.append("?")
                    v.a["alpha"] = al

    # ---------------- inside ----------------
    def estep_in(self):
        """EM_solver.Estep_in_llh (EM_solver.py:91-133)."""
        phi, nu, K = self.phi, self.nu, self.numsites
        for r in self.trees:
            for v in postorder(r):
                d = v.edge_length
                p = math.exp(-d)
                L0 = [0.0] * K
                L1 = [0.0] * K
                al = v.a["alpha"]
                for k in range(K):
                    a = al[k]
                    q = self.Q[k][a] if a not in ("?", "z") else 1.0
                    if v.is_leaf():
                        if a == "?":
                            if self.charMtrx[v.label][k] == "?":
                                t = 1 - (1 - phi) * p ** nu
                            else:
                                t = 1 - p ** nu
                            L0[k] = L1[k] = math.log(t) if t > 0 else MIN_LLH
                        elif a == "z":
                            L0[k] = (nu + 1) * (-d) + math.log(1 - phi) if 1 - phi > 0 else MIN_LLH
                            L1[k] = MIN_LLH
                        else:
                            if (1 - p) * q * (1 - phi) > 0:
                                L0[k] = nu * (-d) + math.log(1 - p) + math.log(q) + math.log(1 - phi)
                            else:
                                L0[k] = MIN_LLH
                            L1[k] = nu * (-d) + math.log(1 - phi) if (1 - phi) > 0 else MIN_LLH
                    else:
                        l0 = 0.0
                        l1 = 0.0
                        for c in v.children:
                            l0 += c.a["L0"][k]
                            l1 += c.a["L1"][k]
                        terms = [l0 + (nu + 1) * (-d)]
                        if a != "z" and q * (1 - p) > 0:
                            terms.append(l1 + math.log(1 - p) + math.log(q) + nu * (-d))
                        if a == "?" and (1 - p ** nu) > 0:
                            terms.append(math.log(1 - p ** nu))
                        L0[k] = log_sum_exp(terms)
                        if a == "z":
                            L1[k] = MIN_LLH
                        elif a != "?" or nu == 0 or p == 1:
                            L1[k] = l1 + nu * (-d)
                        else:
                            L1[k] = log_sum_exp([l1 + nu * (-d), math.log(1 - p ** nu)])
                v.a["L0"] = L0
                v.a["L1"] = L1

    def site_llh(self):
        """Per-site log-likelihood summed over trees (EM_solver.py:135-141)."""
        out = [0.0] * self.numsites
        for r in self.trees:
            for k in range(self.numsites):
                out[k] += r.a["L0"][k]
        return out

    def lineage_llh(self):
        """EM_solver.lineage_llh (EM_solver.py:135-141)."""
        self.estep_in()
        return sum(sum(r.a["L0"]) for r in self.trees)

    def negative_llh(self):
        """ML_solver.negative_llh (ML_solver.py:252-254), EM variant."""
        self.az_partition()
        return -self.lineage_llh()

    # ---------------- outside ----------------
    def _sister(self, v):
        u = v.parent
        for x in u.children:
            if x is not v:
                return x
        return None

    def estep_out(self):
        """EM_solver.Estep_out_llh (EM_solver.py:143-204)."""
        nu, K = self.nu, self.numsites
        for r in self.trees:
            for v in preorder(r):
                d = v.edge_length
                if v.is_root():
                    v.a["A"] = [0.0] * K
                    if nu * d > 0:
                        v.a["X"] = [-nu * d + math.log(1 - math.exp(-d))] * K
                        v.a["out1"] = [math.log(1 - math.exp(-d * nu))] * K
                    else:
                        v.a["X"] = [MIN_LLH] * K
                        v.a["out1"] = [MIN_LLH] * K
                    v.a["oa"] = [dict() for _ in range(K)]
                    v.a["out0"] = [-(1 + nu) * d] * K
                    continue
                u = v.parent
                w = self._sister(v)
                A = [0.0] * K
                X = [0.0] * K
                out0 = [0.0] * K
                out1 = [0.0] * K
                v.a["oa"] = [dict() for _ in range(K)]
                v.a["A"] = A
                v.a["X"] = X
                v.a["out0"] = out0
                v.a["out1"] = out1
                for k in range(K):
                    out0[k] = u.a["out0"][k] + w.a["L0"][k] - (1 + nu) * d
                    A[k] = u.a["out0"][k] + w.a["L0"][k]
                    X[k] = (-nu * d + math.log(1 - math.exp(-d)) + A[k]) if d > 0 else MIN_LLH
                    wa = w.a["alpha"][k]
                    if wa == "z":
                        out1[k] = math.log(1 - math.exp(-d * nu)) + A[k] if nu * d > 0 else MIN_LLH
                    elif wa == "?":
                        X[k] = log_sum_exp([X[k], u.a["X"][k] + w.a["L1"][k] - nu * d])
                        pp = 1 - math.exp(-d * nu)
                        pl = math.log(pp) if pp > 0 else MIN_LLH
                        if nu == 0:
                            out1[k] = u.a["out1"][k]
                        else:
                            out1[k] = log_sum_exp([pl + A[k], pl + u.a["X"][k] + w.a["L1"][k],
                                                   u.a["out1"][k]])
                    else:
                        a0 = wa
                        if a0 not in u.a["oa"][k]:
                            self._out_alpha_up(u, k, a0)
                        B = u.a["oa"][k][a0] + nu * (-d) + w.a["L1"][k]
                        if d > 0 and self.Q[k][a0] > 0:
                            C = A[k] - nu * d + math.log(1 - math.exp(-d)) + math.log(self.Q[k][a0])
                        else:
                            C = MIN_LLH
                        v.a["oa"][k][a0] = log_sum_exp([B, C])
                        X[k] = log_sum_exp([X[k], B])
                        if nu * d > 0:
                            out1[k] = math.log(1 - math.exp(-d * nu)) + log_sum_exp(
                                [A[k], w.a["L1"][k] + u.a["oa"][k][a0]])
                        else:
                            out1[k] = MIN_LLH

    def _out_alpha_up(self, node, k, a0):
        """EM_solver.__out_alpha_up__ (EM_solver.py:206-245)."""
        nu = self.nu
        v = node
        path = []
        while not v.is_root():
            u = v.parent
            w = self._sister(v)
            wa = w.a["alpha"][k]
            if wa != "?" and wa != a0:
                v.a["oa"][k][a0] = (v.a["A"][k] - nu * v.edge_length
                                    + pseudo_log(1 - math.exp(-v.edge_length))
                                    + pseudo_log(self.Q[k][a0]))
                break
            path.append(v)
            v = u
        if v.is_root():
            v.a["oa"][k][a0] = (-nu * v.edge_length + pseudo_log(1 - math.exp(-v.edge_length))
                                + pseudo_log(self.Q[k][a0]))
        while path:
            v = path.pop()
            u = v.parent
            w = self._sister(v)
            B = u.a["oa"][k][a0] + nu * (-v.edge_length) + w.a["L1"][k]
            C = (v.a["A"][k] - nu * v.edge_length + pseudo_log(1 - math.exp(-v.edge_length))
                 + pseudo_log(self.Q[k][a0]))
            v.a["oa"][k][a0] = log_sum_exp([B, C])

    # ---------------- posterior ----------------
    def estep_posterior(self):
        """EM_solver.Estep_posterior (EM_solver.py:247-300)."""
        nu, phi, K = self.nu, self.phi, self.numsites
        for r in self.trees:
            full = r.a["L0"]
            for v in preorder(r):
                d = v.edge_length
                post0 = [0.0] * K
                post1 = [0.0] * K
                S = [[0.0] * K for _ in range(5)]
                for k in range(K):
                    a = v.a["alpha"][k]
                    vin1 = 0.0 if a == "?" else None
                    if v.is_leaf():
                        c = self.charMtrx[v.label][k]
                        if c == 0:
                            vin0 = math.log(1 - phi)
                        else:
                            vin0 = math.log(phi) if (c == "?" and phi > 0) else MIN_LLH
                    else:
                        c1, c2 = v.children
                        vin0 = c1.a["L0"][k] + c2.a["L0"][k]
                    post0[k] = vin0 + v.a["out0"][k] - full[k]
                    post1[k] = vin1 + v.a["out1"][k] - full[k] if vin1 is not None else MIN_LLH
                    if a == "z":
                        S[0][k] = 1.0
                    elif v.is_root():
                        S[0][k] = math.exp(vin0 + (1.0 + nu) * (-d) - v.a["L0"][k])
                        S[2][k] = 0.0 if a != "?" else (1.0 - math.exp(-nu * d)) / math.exp(v.a["L0"][k])
                        S[1][k] = 1.0 - S[0][k] - S[2][k]
                    else:
                        u = v.parent
                        up0 = u.a["post0"][k]
                        up1 = u.a["post1"][k]
                        S[0][k] = math.exp(up0 + vin0 + (1.0 + nu) * (-d) - v.a["L0"][k])
                        if a != "?":
                            S[2][k] = 0.0
                            S[4][k] = 0.0
                        else:
                            S[2][k] = math.exp(up0 - v.a["L0"][k]) * (1.0 - math.exp(-nu * d))
                            t = 1.0 - math.exp(up0) - math.exp(up1)
                            if t == 0 or (1.0 - math.exp(-nu * d)) == 0:
                                S[4][k] = 0.0
                            else:
                                S[4][k] = t * (1.0 - math.exp(-nu * d)) / math.exp(v.a["L1"][k])
                        S[1][k] = math.exp(up0) - S[0][k] - S[2][k]
                        S[3][k] = 1.0 - S[0][k] - S[1][k] - math.exp(post1[k])
                v.a["post0"] = post0
                v.a["post1"] = post1
                v.a["S"] = S

    def estep(self):
        """EM_solver.Estep (EM_solver.py:302-305)."""
        self.estep_in()
        self.estep_out()
        self.estep_posterior()

    # ---------------- M-step ----------------
    def phi_terms(self):
        """R and R-tilde sums of the phi update (EM_solver.py:318-323)."""
        R = 0.0
        Rt = 0.0
        for r in self.trees:
            for v in postorder(r):
                if not v.is_leaf():
                    continue
                row = self.charMtrx[v.label]
                R += sum(1 for x in row if x != "?")
                Rt += sum(1 - math.exp(p) for j, p in enumerate(v.a["post1"]) if row[j] == "?")
        return R, Rt

    def branch_counts(self, local_brlen_opt=True):
        """Clamped, renormalised per-branch S sums (EM_solver.py:328-348)."""
        rows = []
        nodes = []
        for r in self.trees:
            for v in postorder(r):
                if not v.a["poly"] and not (v.a["fixed"] and local_brlen_opt):
                    s = [sum(v.a["S"][j]) for j in range(5)]
                    s = [max(1e-6, x) for x in s]
                    tot = sum(s)
                    s = [x / tot * self.numsites for x in s]
                    rows.append(s)
                    nodes.append(v)
        return rows, nodes

    def ultrametric_constr(self, local_brlen_opt=True):
        """EM_solver.ultrametric_constr (EM_solver.py:53-89): (M, b) over the
        optimised branches in post-order."""
        N = self.num_edges - self.num_polytomy_mark
        if local_brlen_opt:
            for r in self.trees:
                N -= sum(1 for v in postorder(r) if v.a["fixed"])
        constrs = {}
        idx = 0
        for r in self.trees:
            for v in postorder(r):
                if v.is_leaf():
                    v.a["cvec"] = [0.0] * N
                    v.a["const"] = v.edge_length if v.a["fixed"] else 0
                else:
                    c1, c2 = v.children
                    m = tuple(x - y for x, y in zip(c1.a["cvec"], c2.a["cvec"]))
                    mc = tuple(-x for x in m)
                    c = c2.a["const"] - c1.a["const"]
                    if sum(x != 0 for x in m) > 0 and not (m in constrs or mc in constrs):
                        constrs[m] = c
                    v.a["cvec"] = c1.a["cvec"]
                    v.a["const"] = c1.a["const"] + v.edge_length if v.a["fixed"] else c1.a["const"]
                if not v.a["poly"] and not (v.a["fixed"] and local_brlen_opt):
                    v.a["cvec"][idx] = 1
                    idx += 1
        for r in self.trees[1:]:
            r0 = self.trees[0]
            m = tuple(x - y for x, y in zip(r0.a["cvec"], r.a["cvec"]))
            constrs[m] = r.a["const"] - r0.a["const"]
        return [list(m) for m in constrs], [constrs[m] for m in constrs]

    def mstep(self, optimize_phi=True, optimize_nu=True, eps_nu=1e-5, local_brlen_opt=True, ultra=None):
        """EM_solver.Mstep (EM_solver.py:307-459).  ``ultra`` = (M, b) adds the
        ultrametric equality system; that branch problem is solved with scipy's
        SLSQP in branch-length space (an algorithm independent of the device
        solver)."""
        if optimize_phi:
            R, Rt = self.phi_terms()
            phi_star = Rt / (R + Rt)
            if abs(phi_star) < 1 / (self.numsites * len(self.charMtrx)):
                phi_star = 0
        else:
            phi_star = self.phi
        rows, nodes = self.branch_counts(local_brlen_opt)
        nu = self.nu
        use_nu = sum(s[2] for s in rows) > 0 and nu > eps_nu
        use_nu4 = sum(s[4] for s in rows) > 0 and nu > eps_nu
        if ultra is None:
            d_star = [solve_branch(s, nu, use_nu, use_nu4, self.dmin, self.dmax) for s in rows]
        else:
            d_star = solve_branches_constrained(rows, nu, use_nu, use_nu4, self.dmin, self.dmax, ultra[0], ultra[1],
                                                [v.edge_length for v in nodes])
        if optimize_nu:
            nu_star = solve_nu(rows, d_star)
        else:
            nu_star = self.nu
        self.phi = phi_star
        self.nu = nu_star
        for v, d in zip(nodes, d_star):
            v.edge_length = d
        return True, "optimal"

    def em_optimization(self, optimize_phi=True, optimize_nu=True, maxIter=1000, fixed_iters=None,
                        ultra_constr=False):
        """EM_solver.EM_optimization (EM_solver.py:461-512).  ``fixed_iters``
        runs exactly that many iterations with no convergence test."""
        ultra = self.ultrametric_constr() if ultra_constr else None
        pre = self.lineage_llh()
        em_iter = 1
        status = ""
        curr = pre
        self.llh_trace = [pre]
        while em_iter <= (maxIter if fixed_iters is None else fixed_iters):
            self.estep()
            ok, status = self.mstep(optimize_phi=optimize_phi, optimize_nu=optimize_nu, ultra=ultra)
            curr = self.lineage_llh()
            self.llh_trace.append(curr)
            if fixed_iters is None and abs((curr - pre) / pre) < CONV_EPS:
                break
            pre = curr
            em_iter += 1
        return -curr, em_iter, status

    def optimize_one(self, randseed, fixed_phi=None, fixed_nu=None, ultra_constr=False):
        """EM_solver.optimize_one (EM_solver.py:514-527)."""
        rng = _random.Random(randseed)
        x0 = self.ini_all(rng, fixed_phi=fixed_phi, fixed_nu=fixed_nu)
        self.x2params(x0, fixed_phi=fixed_phi, fixed_nu=fixed_nu)
        self.az_partition()
        nllh, it, status = self.em_optimization(optimize_phi=(fixed_phi is None),
                                                optimize_nu=(fixed_nu is None), ultra_constr=ultra_constr)
        self.em_iters = it
        return nllh, status

    def mark_fixed(self, fixed_brlen):
        """Fix branches named by leaf-label anchor pairs at the given lengths
        (ML_solver.py:287-296); all other marks are cleared."""
        for t, r in enumerate(self.trees):
            for v in postorder(r):
                v.a["fixed"] = False
            if fixed_brlen is None:
                continue
            for (a, b), length in fixed_brlen[t].items():
                u = find_lca(r, a, b)
                u.edge_length = length
                u.a["fixed"] = True

    def optimize(self, initials=1, random_seeds=1221, fixed_phi=None, fixed_nu=None, ultra_constr=False,
                 fixed_brlen=None):
        """ML_solver.optimize seed handling (ML_solver.py:256-336) for integer
        seeds and a single round (no retries are needed: the oracle M-step
        never fails)."""
        rng = _random.Random(random_seeds)
        rseeds = [int(rng.random() * 10000) for _ in range(initials)]
        results = []
        for rep in range(initials):
            self.mark_fixed(fixed_brlen)
            nllh, status = self.optimize_one(rseeds[rep], fixed_phi=fixed_phi, fixed_nu=fixed_nu,
                                             ultra_constr=ultra_constr)
            results.append((nllh, rep, self.nu, self.phi, self.branch_lengths(), status))
        results.sort(key=lambda t: (t[0], t[1]))
        return results[0]


# --------------------------------------------------------------------------
# exact solutions of the M-step's concave subproblems (replace MOSEK)
# --------------------------------------------------------------------------
def branch_grad(s, nu, use_nu, use_nu4, d):
    """d/dd of the per-branch objective of EM_solver.py:351-357."""
    S0, S1, S2, S3, S4 = s
    g = -(nu + 1) * S0 - nu * S1 - nu * S3 + S1 / math.expm1(d)
    if use_nu:
        g += S2 * nu / math.expm1(nu * d)
    if use_nu4:
        g += S4 * nu / math.expm1(nu * d)
    return g


def branch_obj(s, nu, use_nu, use_nu4, d):
    S0, S1, S2, S3, S4 = s
    f = -(nu + 1) * S0 * d - nu * S1 * d + S1 * math.log(-math.expm1(-d)) - nu * S3 * d
    if use_nu:
        f += S2 * math.log(-math.expm1(-nu * d))
    if use_nu4:
        f += S4 * math.log(-math.expm1(-nu * d))
    return f


def solve_branch(s, nu, use_nu, use_nu4, dmin, dmax):
    """argmax over [dmin, dmax] of the strictly concave branch objective, by
    bisection on its derivative to full double precision."""
    if branch_grad(s, nu, use_nu, use_nu4, dmin) <= 0:
        return dmin
    if branch_grad(s, nu, use_nu, use_nu4, dmax) >= 0:
        return dmax
    lo, hi = dmin, dmax
    while True:
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            return mid
        if branch_grad(s, nu, use_nu, use_nu4, mid) > 0:
            lo = mid
        else:
            hi = mid


def solve_branches_constrained(rows, nu, use_nu, use_nu4, dmin, dmax, M, b, d0):
    """argmax of the summed branch objectives subject to M d = b and the box,
    with scipy SLSQP (analytic gradient), from a feasible-ish start."""
    import numpy as np
    from scipy.optimize import minimize

    S = np.asarray(rows, dtype=float)
    A = (nu + 1) * S[:, 0] + nu * S[:, 1] + nu * S[:, 3]
    S1 = S[:, 1]
    B = (S[:, 2] if use_nu else 0.0) + (S[:, 4] if use_nu4 else 0.0)

    def f(d):
        val = -A @ d + S1 @ np.log(-np.expm1(-d))
        if nu > 0:
            val = val + B @ np.log(-np.expm1(-nu * d)) if np.any(B) else val
        return -val

    def g(d):
        gr = -A + S1 / np.expm1(d)
        if nu > 0 and np.any(B):
            gr = gr + B * nu / np.expm1(nu * d)
        return -gr

    Mm = np.asarray(M, dtype=float).reshape(-1, len(rows))
    bb = np.asarray(b, dtype=float)
    cons = [{"type": "eq", "fun": lambda d: Mm @ d - bb, "jac": lambda d: Mm}] if len(bb) else []
    x0 = np.clip(np.asarray(d0, dtype=float), dmin * 1.01, dmax * 0.99)
    res = minimize(f, x0, jac=g, method="SLSQP", bounds=[(dmin, dmax)] * len(rows), constraints=cons,
                   options={"ftol": 1e-16, "maxiter": 5000})
    return list(res.x)


def nu_grad(rows, d, nu):
    """d/dnu of the nu objective of EM_solver.py:392-403."""
    use2 = sum(s[2] for s in rows) > 0
    use4 = sum(s[4] for s in rows) > 0
    g = 0.0
    for s, dv in zip(rows, d):
        g -= dv * (s[0] + s[1] + s[3])
        if use2:
            g += s[2] * dv / math.expm1(nu * dv)
        if use4:
            g += s[4] * dv / math.expm1(nu * dv)
    return g


def nu_obj(rows, d, nu):
    use2 = sum(s[2] for s in rows) > 0
    use4 = sum(s[4] for s in rows) > 0
    f = 0.0
    for s, dv in zip(rows, d):
        f += -(nu + 1) * s[0] * dv - nu * s[1] * dv + s[1] * math.log(-math.expm1(-dv)) - nu * s[3] * dv
        if use2:
            f += s[2] * math.log(-math.expm1(-nu * dv))
        if use4:
            f += s[4] * math.log(-math.expm1(-nu * dv))
    return f


def solve_nu(rows, d):
    """argmax over [0, 1] of the concave nu objective (bisection)."""
    if nu_grad(rows, d, 1.0) >= 0:
        return 1.0
    lo, hi = 0.0, 1.0
    while True:
        mid = 0.5 * (lo + hi)
        if mid <= lo or mid >= hi:
            return mid
        if nu_grad(rows, d, mid) > 0:
            lo = mid
        else:
            hi = mid


# --------------------------------------------------------------------------
# input readers (restating laml_libs/sequence_lib.py:102-175 for test data)
# --------------------------------------------------------------------------
_RECOGNIZED_MISSING = {"-", "?", "-1"}


def read_charmtrx(path, delimiter=",", masked_symbol=None):
    """sequence_lib.read_charMtrx (sequence_lib.py:139-175): missing symbols
    become '?', everything else int."""
    D = {}
    with open(path) as f:
        head = f.readline().strip().split(delimiter)
        seen = {masked_symbol} if masked_symbol is not None else set()
        for line in f:
            parts = line.strip().split(delimiter)
            if not parts or parts == [""]:
                continue
            seq = []
            for x in parts[1:]:
                miss = (x in seen or x in _RECOGNIZED_MISSING or x.isalpha())
                if not miss:
                    try:
                        miss = int(x) < 0
                    except ValueError:
                        miss = False
                if miss:
                    seen.add(x)
                    seq.append("?")
                else:
                    seq.append(int(x))
            D[parts[0]] = seq
    return D


def read_priors_csv(path, k):
    """Prior file reader as in unit_tests_EMSolver.py:542-556."""
    Q = [{0: 0} for _ in range(k)]
    seen = set()
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            si, st, pr = line.split(",")
            si = int(si)
            if si not in seen:
                seen.add(si)
            Q[len(seen) - 1][int(st)] = float(pr)
    return Q
