"""CPU oracle for LAML's ML_solver likelihood (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.

ML_solver evaluates the same PMM likelihood as EM_solver but on the tree as
given, polytomies included, and with its own leaf rules.  This restates, in
scalar FP64 log space:

  prior normalisation     laml_libs/ML_solver.py:69-74
  az_partition            laml_libs/ML_solver.py:125-154
                          (a leaf entry -1 is labelled '?', like '?')
  lineage_llh             laml_libs/ML_solver.py:156-193
                          (an internal node sums L0/L1 over all its children;
                          a '?' leaf has L0 = L1 = log(1 - (1-phi) p^nu))
  negative_llh            laml_libs/ML_solver.py:252-254
  ultrametric_constr      laml_libs/ML_solver.py:95-112
  optimize_one (SLSQP)    laml_libs/ML_solver.py:338-377

``lineage_llh_mp`` evaluates the same recursion in mpmath at 40 digits: where
nu * d is tiny the reference's double-precision 1 - p**nu cancels (relative
error about 1e-16 / (nu d)), and the device, which forms it with expm1, is
compared with the exact value instead.

Pinned in tests by the reference's ML known-answer tests
(laml_unit_tests/unit_tests_MLSolver.py), by the reference's legacy
compute-likelihood case on a tree with a four-way polytomy
(case_tests/old/test6), and by the dense per-state Felsenstein oracle on
random multifurcating trees.
"""
import math

from .em_oracle import MIN_LLH, parse_newick, postorder, suppress_unifurcations

EPS = 1e-10
DMIN = 0.005
DMAX = 10.0


def normalise_Q(Q):
    out = []
    for q in Q:
        s = sum(q[x] for x in q)
        out.append({x: q[x] / s for x in q})
    return out


def az_partition(nodes, charMtrx, K):
    """node.a['alpha'][k] in {'z', '?', state} for post-order ``nodes``."""
    for v in nodes:
        if v.is_leaf():
            row = charMtrx[v.label]
            lab = []
            for k in range(K):
                x = row[k]
                if x == 0:
                    lab.append('z')
                elif x == -1 or x == '?':
                    lab.append('?')
                else:
                    lab.append(x)
            v.a['alpha'] = lab
        else:
            lab = []
            for k in range(K):
                S = set(c.a['alpha'][k] for c in v.children)
                R = S - {'z', '?', -1}
                if 'z' in S or len(R) > 1:
                    lab.append('z')
                elif len(R) == 1:
                    lab.append(next(iter(R)))
                else:
                    lab.append('?')
            v.a['alpha'] = lab


def lineage_llh(nodes, Q, nu, phi, K):
    """Sum over sites of the log-likelihood of post-order ``nodes`` (each
    with 'alpha' from az_partition and a numeric edge_length)."""
    llh = [0.0] * K
    for v in nodes:
        d = v.edge_length
        p = math.exp(-d)
        L0 = [0.0] * K
        L1 = [0.0] * K
        alpha = v.a['alpha']
        for k in range(K):
            a = alpha[k]
            if a != 'z':
                q = Q[k][a] if a != '?' else 1.0
                if v.is_leaf():
                    if a == '?':
                        m = 1.0 - (1.0 - phi) * p ** nu
                        L0[k] = L1[k] = math.log(m) if m > 0 else MIN_LLH
                    else:
                        L0[k] = (nu * (-d) + math.log(1 - p) + math.log(q) + math.log(1 - phi)
                                 if (1 - p) * q * (1 - phi) > 0 else MIN_LLH)
                        L1[k] = nu * (-d) + math.log(1 - phi)
                else:
                    l0 = sum(c.a['L0'][k] for c in v.children)
                    l1 = sum(c.a['L1'][k] for c in v.children)
                    extra = (1.0 - p ** nu) if a == '?' else 0.0
                    x0 = math.exp(l0 + (nu + 1) * (-d)) + q * (1 - p) * math.exp(l1 + nu * (-d)) + extra
                    x1 = math.exp(l1 + nu * (-d)) + extra
                    L0[k] = MIN_LLH if x0 == 0 else math.log(x0)
                    L1[k] = MIN_LLH if x1 == 0 else math.log(x1)
                if v.is_root() or v.parent.a['alpha'][k] == 'z':
                    llh[k] += L0[k]
            else:
                llh[k] += -d * (1 + nu) + (math.log(1 - phi) if v.is_leaf() else 0.0)
        v.a['L0'] = L0
        v.a['L1'] = L1
    return sum(llh)


def lineage_llh_mp(nodes, Q, nu, phi, K, dps=40):
    """lineage_llh in mpmath arithmetic at ``dps`` digits (linear space, no
    cancellation), for parameters where the reference's double-precision
    terms such as 1 - p**nu lose digits (nu * d below about 1e-6)."""
    import mpmath

    with mpmath.workdps(dps):
        nu_m, phi_m = mpmath.mpf(nu), mpmath.mpf(phi)
        llh = mpmath.mpf(0)
        L0s, L1s = {}, {}
        for v in nodes:
            d = mpmath.mpf(v.edge_length)
            p = mpmath.exp(-d)
            pn = p ** nu_m
            L0, L1 = [None] * K, [None] * K
            for k in range(K):
                a = v.a['alpha'][k]
                if a != 'z':
                    q = mpmath.mpf(Q[k][a]) if a != '?' else mpmath.mpf(1)
                    if v.is_leaf():
                        if a == '?':
                            L0[k] = L1[k] = 1 - (1 - phi_m) * pn
                        else:
                            L0[k] = pn * (1 - p) * q * (1 - phi_m)
                            L1[k] = pn * (1 - phi_m)
                    else:
                        l0 = mpmath.mpf(1)
                        l1 = mpmath.mpf(1)
                        for c in v.children:
                            l0 *= L0s[id(c)][k]
                            l1 *= L1s[id(c)][k]
                        extra = (1 - pn) if a == '?' else 0
                        L0[k] = l0 * pn * p + q * (1 - p) * l1 * pn + extra
                        L1[k] = l1 * pn + extra
                    if v.is_root() or v.parent.a['alpha'][k] == 'z':
                        llh += mpmath.log(L0[k])
                else:
                    llh += -d * (1 + nu_m) + (mpmath.log(1 - phi_m) if v.is_leaf() else 0)
            L0s[id(v)], L1s[id(v)] = L0, L1
        return float(llh)


def negative_llh_mp(tree_strs, charMtrx, Q, nu, phi, dps=40):
    """negative_llh evaluated by lineage_llh_mp."""
    if isinstance(tree_strs, str):
        tree_strs = [tree_strs]
    Qn = normalise_Q(Q)
    K = len(next(iter(charMtrx.values())))
    _, posts = read_trees(tree_strs)
    total = 0.0
    for nodes in posts:
        az_partition(nodes, charMtrx, K)
        total += lineage_llh_mp(nodes, Qn, nu, phi, K, dps)
    return -total


def read_trees(tree_strs):
    """Parsed trees (unifurcations suppressed) and their post-order lists."""
    roots = [suppress_unifurcations(parse_newick(t)) for t in tree_strs]
    return roots, [postorder(r) for r in roots]


def negative_llh(tree_strs, charMtrx, Q, nu, phi):
    """ML_solver(trees, {'charMtrx'}, {'Q'}, {'nu', 'phi'}).negative_llh()."""
    if isinstance(tree_strs, str):
        tree_strs = [tree_strs]
    Qn = normalise_Q(Q)
    K = len(next(iter(charMtrx.values())))
    _, posts = read_trees(tree_strs)
    total = 0.0
    for nodes in posts:
        az_partition(nodes, charMtrx, K)
        total += lineage_llh(nodes, Qn, nu, phi, K)
    return -total


def ultrametric_constr(posts, N):
    """ML_solver.ultrametric_constr: rows of M with M x = 0 (binary trees)."""
    M = []
    idx = 0
    cons = {}
    for nodes in posts:
        for v in nodes:
            if v.is_leaf():
                c = [0.0] * N
            else:
                c1, c2 = v.children
                M.append([x - y for x, y in zip(cons[id(c1)], cons[id(c2)])])
                c = cons[id(c1)]
            c[idx] = 1
            cons[id(v)] = c
            idx += 1
    roots = [nodes[-1] for nodes in posts]
    for r in roots[1:]:
        M.append([x - y for x, y in zip(cons[id(roots[0])], cons[id(r)])])
    return M


def slsqp_optimize(tree_strs, charMtrx, Q, x0, fixed_nu=None, fixed_phi=None, fixed_idx=(), ultra=False,
                   dmin=DMIN, dmax=DMAX):
    """ML_solver.optimize_one's SciPy call (finite-difference gradients) from
    the start point x0 = [branch lengths in post-order..., nu, phi];
    ``fixed_idx`` lists branch positions held at their x0 value.  Returns
    (nllh, x, success)."""
    import numpy as np
    from scipy import optimize
    from scipy.sparse import csr_matrix

    if isinstance(tree_strs, str):
        tree_strs = [tree_strs]
    Qn = normalise_Q(Q)
    K = len(next(iter(charMtrx.values())))
    _, posts = read_trees(tree_strs)
    for nodes in posts:
        az_partition(nodes, charMtrx, K)
    ne = sum(len(n) for n in posts)

    def fun(x):
        i = 0
        for nodes in posts:
            for v in nodes:
                v.edge_length = x[i]
                i += 1
        nu = x[ne] if fixed_nu is None else fixed_nu
        phi = x[ne + 1] if fixed_phi is None else fixed_phi
        return -sum(lineage_llh(nodes, Qn, nu, phi, K) for nodes in posts)

    lo = [dmin] * ne + [EPS if fixed_nu is None else fixed_nu - EPS, EPS if fixed_phi is None else fixed_phi - EPS]
    hi = [dmax] * ne + [10 if fixed_nu is None else fixed_nu + EPS, 0.99 if fixed_phi is None else fixed_phi + EPS]
    cons = []
    if fixed_idx:
        A = np.zeros((len(fixed_idx), len(x0)))
        b = []
        for r, i in enumerate(fixed_idx):
            A[r, i] = 1
            b.append(x0[i])
        cons.append(optimize.LinearConstraint(csr_matrix(A), b, b))
    if ultra:
        M = ultrametric_constr(posts, len(x0))
        cons.append(optimize.LinearConstraint(csr_matrix(M), [0] * len(M), [0] * len(M)))
    out = optimize.minimize(fun, x0, method="SLSQP", options={'maxiter': 1000},
                            bounds=optimize.Bounds(lo, hi), constraints=cons)
    return float(out.fun), out.x, bool(out.success)
