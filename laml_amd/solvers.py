"""Drop-in solver classes with the surface of LAML's solvers.

``EM_solver`` mirrors laml_libs/EM_solver.py (constructor, negative_llh,
Estep, Mstep, EM_optimization, optimize_one, optimize, score_tree,
get_tree_newick, get_params).  Tree handling, seeding and the optimisation
driver stay on the host exactly as in the reference; every likelihood, E-step
and M-step evaluation runs in the HIP kernels of ``liblaml_pmm.so`` on the GPU.
There is no CPU fallback: constructing the device forest without the library
or without a GPU raises.

``fastEM_solver`` keeps the reference's fastEM constructor/score_tree surface
but evaluates with the same exact FP64 engine (the reference's fastEM is an
FP32 approximation with a different M-step, laml/fast_em/laml.py:62-110).
"""
import random as _random
import warnings
from copy import deepcopy
from math import isfinite

import numpy as np
import torch

from ._lib import LAML_MSTEP_FAILED_D, LAML_MSTEP_FAILED_NU, LAML_OPT_NU, LAML_OPT_PHI
from .forest import DMAX, DMIN, Forest, ProblemSpec, SiteTable, resolved_tree_arrays, tree_arrays_from_tree
from .lca import find_LCAs
from .tree import read_tree_newick
from .ultra import HeightModel, UltraInfeasible, ultrametric_constr as _ultrametric_constr

# laml_libs/core_constants.py:3-10
MIN_LLH = -float("inf")
EPS = 1e-10
CONV_EPS = 1e-8
NNI_CONV_EPS = 1e-15
CHKPT_FREQ = 10
DEFAULT_STRATEGY = {'resolve_search_only': False, 'only_marked': False, 'ultra_constr': False,
                    'fixed_phi': None, 'fixed_nu': None, 'local_brlen_opt': True,
                    'nodes_to_recompute': None}



def mstep_status(bits):
    """(success, status) of EM_solver.Mstep from the device's M-step status
    bits (EM_solver.py:450-459: "failed_d" when the branch solve failed, and
    ",failed_nu" -- replacing it -- when the nu solve failed)."""
    bits = int(bits)
    if bits == 0:
        return True, "optimal"
    status = ""
    if bits & LAML_MSTEP_FAILED_D:
        status += "failed_d"
    if bits & LAML_MSTEP_FAILED_NU:
        status = ",failed_nu"
    return False, status

class Params:
    def __init__(self, nu, phi):
        self.nu = nu
        self.phi = phi


def _device_index():
    if not torch.cuda.is_available():
        raise RuntimeError("laml_amd solvers need a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.cuda.current_device()


class EM_solver:
    """GPU-backed replacement of laml_libs.EM_solver.EM_solver."""
    solver_name = "EM_solver"

    def __init__(self, treeList, data, prior, params={'nu': 0, 'phi': 0, 'sigma': 0}):
        self._read_inputs(treeList, data, prior, params)
        self.solver_name = "EM_solver"
        # EM_solver.__init__ / __mark_polytomies__ (EM_solver.py:19-43)
        self.has_polytomy = False
        self.__mark_polytomies__(eps_len=self.dmin * 0.01)
        self.num_edges = sum(sum(1 for _ in t.traverse_postorder()) for t in self.trees)
        for t in self.trees:
            for v in t.traverse_postorder():
                if not hasattr(v, 'mark_fixed'):
                    v.mark_fixed = False

    def _read_inputs(self, treeList, data, prior, params):
        """ML_solver.__init__ (ML_solver.py:19-86): trees (unifurcations
        suppressed), character matrix (uninformative sites dropped when that
        reconciles it with the priors), normalised priors, parameters."""
        charMtrx = data['charMtrx']
        Q = prior['Q']
        self.trees = []
        self.num_edges = 0
        for tree in treeList:
            t = read_tree_newick(tree) if isinstance(tree, str) else tree
            t.suppress_unifurcations()
            self.num_edges += sum(1 for _ in t.traverse_postorder())
            self.trees.append(t)
        site_matrix = list(zip(*charMtrx.values()))
        uninformative = [i for i, site in enumerate(site_matrix) if all(v in (0, -1, '?') for v in site)]
        if len(Q) != len(site_matrix):
            print(f"Warning: Q has {len(Q)} entries, {len(site_matrix)} sites in charMtrx. "
                  f"Identified {len(uninformative)} uninformative sites.")
            if len(uninformative) == max(len(Q), len(site_matrix)) - min(len(Q), len(site_matrix)):
                keep = [i for i in range(len(site_matrix)) if i not in set(uninformative)]
                charMtrx = {c: [v[i] for i in keep] for c, v in charMtrx.items()}
            else:
                print("Warning: Removing the uninformative sites would not help us resolve the discrepancy "
                      "between the priors and the character matrix, so let's not change anything.")
        self.charMtrx = charMtrx
        self.Q = []
        for Qi in Q:
            s = sum(Qi[x] for x in Qi)
            self.Q.append({x: Qi[x] / s for x in Qi})
        self.params = Params(params['nu'], params['phi'])
        self.numsites = len(self.charMtrx[next(iter(self.charMtrx.keys()))])
        self.dmin = DMIN
        self.dmax = DMAX
        self._forest = None
        self._forest_key = None
        self._sites = None

    # ------------------------------------------------------------------
    def __mark_polytomies__(self, eps_len=0):
        self.has_polytomy = False
        self.num_polytomy_mark = 0
        for tree in self.trees:
            for node in tree.traverse_preorder():
                node.polytomy_mark = False
                if len(node.children) > 2:
                    self.has_polytomy = True
            for nn in tree.resolve_polytomies():
                nn.polytomy_mark = True
                nn.edge_length = eps_len
                self.num_polytomy_mark += 1
                self.has_polytomy = True

    def get_tree_newick(self):
        return [tree.newick() for tree in self.trees]

    def get_params(self):
        return {'phi': self.params.phi, 'nu': self.params.nu}

    # ---------------- device forest ----------------
    def _site_table(self):
        if self._sites is None:
            self._sites = SiteTable(self.Q)
        return self._sites

    def _topology_key(self):
        key = []
        for t in self.trees:
            key.append(tuple((id(v), id(v.parent), bool(v.polytomy_mark), bool(getattr(v, 'mark_fixed', False)))
                             for v in t.traverse_postorder()))
        return tuple(key)

    def ultrametric_constr(self, local_brlen_opt=True):
        """EM_solver.ultrametric_constr (EM_solver.py:53-89): (M, b)."""
        return _ultrametric_constr(self.trees, local_brlen_opt=local_brlen_opt)

    def _enable_ultra(self, F, d):
        """Height-class model of the current tree/fixed set on the device, with
        strictly feasible start heights from the branch lengths d (level
        order).  Returns False when the constraints are infeasible."""
        key = self._forest_key
        try:
            if getattr(self, "_ultra_key", None) != key:
                self._ultra_model = HeightModel(F, d)
                self._ultra_key = key
                F.set_ultrametric(self._ultra_model)
            t0 = self._ultra_model.start_heights(d, F.free, self.dmin, self.dmax)
        except UltraInfeasible:
            return False
        F.set_heights(t0)
        return True

    def _disable_ultra(self, F):
        if getattr(self, "_ultra_key", None) is not None:
            F.set_ultrametric(None)
            self._ultra_key = None

    def _leaf_row(self, label):
        """The character row a leaf is encoded from (EM_solver reads -1 as
        silenced and '?' as missing, EM_solver.py:104-109)."""
        return self.charMtrx[label]

    def _tree_arrays(self, tree, codes_of, free_of, poly_of):
        """Device arrays of one tree and, per node of the tree in post-order,
        its index in them."""
        ta = tree_arrays_from_tree(tree, codes_of, free_of, poly_of)
        return ta, np.arange(ta.n_nodes)

    def _node_gids(self, F, ti):
        """Level-order device index of each node of tree ti (post-order)."""
        return F.inv[F.tree_base[ti] + self._post_index[ti]]

    def _problem_spec(self, flags):
        sites = self._site_table()
        rows = {}

        def codes_of(v):
            if v.label not in self.charMtrx:
                raise KeyError("leaf %r is not in the character matrix" % (v.label,))
            r = rows.get(v.label)
            if r is None:
                r = sites.encode_row(self._leaf_row(v.label))
                rows[v.label] = r
            return r

        def free_of(v):
            return not v.polytomy_mark and not getattr(v, 'mark_fixed', False)

        def poly_of(v):
            return bool(v.polytomy_mark)

        built = [self._tree_arrays(t, codes_of, free_of, poly_of) for t in self.trees]
        arrays = [a for a, _ in built]
        self._post_index = [ix for _, ix in built]
        nobs = 0
        for t in self.trees:
            for v in t.traverse_leaves():
                nobs += sum(1 for x in self.charMtrx[v.label] if x != '?')
        phimin = 1.0 / (self.numsites * len(self.charMtrx))
        return ProblemSpec(arrays, sites, nobs, phimin, flags)

    def _forest_for(self, flags=LAML_OPT_PHI | LAML_OPT_NU):
        key = self._topology_key()
        if self._forest is None or self._forest_key != key:
            dev = _device_index()
            if self._forest is not None:
                self._forest.close()
            spec = self._problem_spec(flags)
            self._forest = Forest([spec], device=dev, dmin=self.dmin, dmax=self.dmax)
            self._forest_key = key
            self._forest_flags = flags
            self._ultra_key = None
        elif self._forest_flags != flags:
            self._forest.set_flags([flags])
            self._forest_flags = flags
        return self._forest

    def _device_params(self, F):
        lengths = []
        for ti, t in enumerate(self.trees):
            L = np.zeros(F.tree_base[ti + 1] - F.tree_base[ti])
            L[self._post_index[ti]] = [0.0 if v.edge_length is None else float(v.edge_length)
                                       for v in t.traverse_postorder()]
            lengths.append(L)
        d = torch.tensor(F.lengths_to_level(lengths), dtype=torch.float64, device=F.device)
        nu = torch.tensor([float(self.params.nu)], dtype=torch.float64, device=F.device)
        phi = torch.tensor([float(self.params.phi)], dtype=torch.float64, device=F.device)
        return d, nu, phi

    def _store_params(self, F, d, nu, phi, only_free=True):
        per_tree = F.level_to_tree_lengths(d.cpu().numpy())
        for t, L, ix in zip(self.trees, per_tree, self._post_index):
            for v, x in zip(t.traverse_postorder(), L[ix]):
                if only_free and (v.polytomy_mark or getattr(v, 'mark_fixed', False)):
                    continue
                v.edge_length = float(x)
        self.params.nu = float(nu.item())
        self.params.phi = float(phi.item())

    # ---------------- likelihood ----------------
    def az_partition(self):
        """Builds the device forest; az-partition labels are computed on the
        device (ML_solver.py:125-154) and can be read with node_labels()."""
        self._forest_for()

    def node_labels(self):
        """Per-tree dict node -> list of labels ('z', '?', or the edit state),
        as node.alpha of the reference's az_partition."""
        F = self._forest_for()
        leaf_l, int_l = F.labels()
        sites = self._site_table()
        inv_maps = [{j: s for s, j in m.items()} for m in sites.maps]
        K = self.numsites
        out = []
        nl = F.num_leaves
        for ti, t in enumerate(self.trees):
            res = {}
            gids = self._node_gids(F, ti)
            for pi, v in enumerate(t.traverse_postorder()):
                g = gids[pi]
                row = leaf_l[g * K:(g + 1) * K] if g < nl else int_l[(g - nl) * K:(g - nl + 1) * K]
                res[v] = ['z' if x == 0 else ('?' if x < 0 else inv_maps[k][int(x)]) for k, x in enumerate(row)]
            out.append(res)
        return out

    def lineage_llh(self):
        F = self._forest_for()
        d, nu, phi = self._device_params(F)
        return -float(F.nll(d, nu, phi).item())

    def __llh__(self):
        return self.lineage_llh()

    def negative_llh(self):
        """ML_solver.negative_llh (ML_solver.py:252-254)."""
        return -self.lineage_llh()

    def site_llh(self):
        """Per-site log-likelihood of each tree, list of arrays [K]."""
        F = self._forest_for()
        d, nu, phi = self._device_params(F)
        F.nll(d, nu, phi, site_llh=True)
        a = F.site_llh.cpu().numpy()
        return [a[F.site_base[i]:F.site_base[i + 1]] for i in range(F.num_trees)]

    # ---------------- EM ----------------
    def Estep(self):
        """EM_solver.Estep (EM_solver.py:302-305) on the device.  Stores the
        per-branch sums over sites of S0..S4 in node.S_sum and returns
        (nllh, R_tilde)."""
        F = self._forest_for()
        d, nu, phi = self._device_params(F)
        counts, rt, nll = F.estep(d, nu, phi)
        c = counts.view(-1, 5).cpu().numpy()
        for ti, t in enumerate(self.trees):
            gids = self._node_gids(F, ti)
            for pi, v in enumerate(t.traverse_postorder()):
                v.S_sum = c[gids[pi]].tolist()
        self._last_estep = (d, nu, phi)
        return float(nll.item()), float(rt.item())

    def Mstep(self, optimize_phi=True, optimize_nu=True, verbose=1, eps_nu=1e-5, eps_s=1e-6,
              ultra_constr_cache=None, local_brlen_opt=True):
        """EM_solver.Mstep (EM_solver.py:307-459) on the device, after Estep.
        A non-None ultra_constr_cache selects the ultrametric-constrained
        branch problem (the height-class model replaces the (M, b) matrix)."""
        flags = (LAML_OPT_PHI if optimize_phi else 0) | (LAML_OPT_NU if optimize_nu else 0)
        F = self._forest_for(flags)
        d, nu, phi = self._device_params(F)
        if ultra_constr_cache is not None:
            if not self._enable_ultra(F, d.cpu().numpy()):
                return False, "d_infeasible"
        else:
            self._disable_ultra(F)
        F.mstep(d, nu, phi)
        self._store_params(F, d, nu, phi)
        return mstep_status(F.mstep_status()[0])

    def EM_optimization(self, verbose=1, optimize_phi=True, optimize_nu=True, ultra_constr=False, maxIter=1000):
        """EM_solver.EM_optimization (EM_solver.py:461-512), whole loop on
        the device.  Returns (nllh, em_iter, status)."""
        flags = (LAML_OPT_PHI if optimize_phi else 0) | (LAML_OPT_NU if optimize_nu else 0)
        F = self._forest_for(flags)
        d, nu, phi = self._device_params(F)
        if ultra_constr:
            if not self._enable_ultra(F, d.cpu().numpy()):
                if verbose >= 0:
                    print("Warning: EM failed to optimize parameters in one Mstep due to infeasible constraints")
                return float(F.nll(d, nu, phi).item()), 1, "d_infeasible"
        else:
            self._disable_ultra(F)
        if verbose >= 0:
            pre = F.nll(d, nu, phi).item()
            print("Initial phi: " + str(self.params.phi) + ". Initial nu: " + str(self.params.nu)
                  + ". Initial nllh: " + str(pre))
        nll, iters = F.em(d, nu, phi, maxIter, CONV_EPS)
        self._store_params(F, d, nu, phi)
        nllh = float(nll.item())
        em_iter = int(iters.item())
        # the status of the last M-step, as the reference's loop returns it
        ok, status = mstep_status(F.mstep_status()[0])
        if not ok and verbose >= 0:
            print("Warning: EM failed to optimize parameters in one Mstep.")
        if em_iter > maxIter and verbose >= 0:
            print("Warning: exceeded maximum number of EM iterations (" + str(maxIter) + " iters)!")
        return nllh, em_iter, status

    # ---------------- initial points (ML_solver.py:195-247) ----------------
    def ini_brlens(self):
        x = [_random.random() * (self.dmax / 2 - 2 * self.dmin) + 2 * self.dmin for _ in range(self.num_edges)]
        idx = 0
        for tree in self.trees:
            for node in tree.traverse_postorder():
                if node.edge_length is not None:
                    x[idx] = max(2 * self.dmin, node.edge_length)
                idx += 1
        return x

    def ini_nu(self, fixed_nu=None):
        return _random.random() * 0.99 if fixed_nu is None else fixed_nu

    def ini_phi(self, fixed_phi=None):
        return _random.random() * 0.99 if fixed_phi is None else fixed_phi

    def ini_all(self, fixed_phi=None, fixed_nu=None):
        return self.ini_brlens() + [self.ini_nu(fixed_nu=fixed_nu), self.ini_phi(fixed_phi=fixed_phi)]

    def x2brlen(self, x):
        i = 0
        for tree in self.trees:
            for node in tree.traverse_postorder():
                if not node.polytomy_mark:
                    node.edge_length = x[i]
                    i += 1

    def x2nu(self, x, fixed_nu=None):
        self.params.nu = x[self.num_edges] if fixed_nu is None else fixed_nu

    def x2phi(self, x, fixed_phi=None):
        self.params.phi = x[self.num_edges + 1] if fixed_phi is None else fixed_phi

    def x2params(self, x, fixed_nu=None, fixed_phi=None):
        self.x2brlen(x)
        self.x2nu(x, fixed_nu=fixed_nu)
        self.x2phi(x, fixed_phi=fixed_phi)

    # ---------------- drivers ----------------
    def optimize_one(self, randseed, fixed_phi=None, fixed_nu=None, verbose=1, ultra_constr=False):
        """EM_solver.optimize_one (EM_solver.py:514-527)."""
        _random.seed(a=randseed)
        x0 = self.ini_all(fixed_phi=fixed_phi, fixed_nu=fixed_nu)
        self.x2params(x0, fixed_phi=fixed_phi, fixed_nu=fixed_nu)
        self.az_partition()
        nllh, em_iter, status = self.EM_optimization(verbose=verbose, optimize_phi=(fixed_phi is None),
                                                     optimize_nu=(fixed_nu is None), ultra_constr=ultra_constr)
        if verbose >= 0:
            print("EM finished after " + str(em_iter) + " iterations.")
            print("Optimal phi: " + str(self.params.phi) + ". Optimal nu: " + str(self.params.nu)
                  + ". Optimal nllh: " + str(nllh))
        self.em_iterations = em_iter
        return nllh, status

    def optimize(self, initials=20, fixed_phi=None, fixed_nu=None, fixed_brlen=None, verbose=1, max_trials=100,
                 random_seeds=None, ultra_constr=False):
        """ML_solver.optimize (ML_solver.py:256-336)."""
        results = []
        all_failed = True
        all_trials = 0
        if random_seeds is None:
            rseeds = [int(_random.random() * 10000) for _ in range(initials)]
        elif type(random_seeds) == int:
            if verbose >= 0:
                print("Global random seed: " + str(random_seeds))
            _random.seed(a=random_seeds)
            rseeds = [int(_random.random() * 10000) for _ in range(initials)]
        elif type(random_seeds) == list:
            if len(random_seeds) < initials:
                if verbose >= 0:
                    print("Fatal: the number of random seeds is smaller than the number of initials!")
                return None
            rseeds = random_seeds[:initials]
        else:
            if verbose >= 0:
                print("Fatal: incorrect random_seeds type provided")
            return None
        while all_failed and all_trials < max_trials:
            for rep in range(initials):
                randseed = rseeds[rep] + all_trials
                if verbose >= 0:
                    print("Initial point " + str(rep + 1) + ". Random seed: " + str(randseed))
                for t, tree in enumerate(self.trees):
                    for node in tree.traverse_postorder():
                        node.mark_fixed = False
                    if fixed_brlen is None:
                        continue
                    fixed_nodes = find_LCAs(tree, list(fixed_brlen[t].keys()))
                    for i, (a, b) in enumerate(fixed_brlen[t]):
                        u = fixed_nodes[i]
                        u.edge_length = fixed_brlen[t][(a, b)]
                        u.mark_fixed = True
                nllh, status = self.optimize_one(randseed, fixed_phi=fixed_phi, fixed_nu=fixed_nu, verbose=verbose,
                                                 ultra_constr=ultra_constr)
                if nllh is not None and isfinite(nllh):
                    all_failed = False
                    processed = []
                    for tree in self.trees:
                        tc = read_tree_newick(tree.newick())
                        tc.collapse_short_branches(self.dmin * 0.01)
                        processed.append(tc.newick())
                    results.append((nllh, rep, deepcopy(self.params), processed, status))
                elif verbose >= 0:
                    print("Fatal: failed to optimize using initial point " + str(rep + 1))
            all_trials += initials
        if all_failed:
            if verbose >= 0:
                print("Fatal: Optimization failed on more than " + str(max_trials) + " retries. Aborting ...")
            return None
        results.sort(key=lambda r: (r[0], r[1]))
        best_nllh, _, best_params, best_trees, status = results[0]
        self.trees = [read_tree_newick(t) for t in best_trees]
        for t in self.trees:
            for v in t.traverse_postorder():
                v.polytomy_mark = False
                v.mark_fixed = False
        self.params = best_params
        self._forest = None
        if verbose >= 0:
            print("Numerical optimization finished successfully")
        return best_nllh, status

    def score_tree(self, strategy={'ultra_constr': False, 'fixed_phi': None, 'fixed_nu': None, 'fixed_brlen': None}):
        """ML_solver.score_tree (ML_solver.py:114-123)."""
        nllh, status = self.optimize(initials=1, verbose=-1, ultra_constr=strategy['ultra_constr'],
                                     fixed_phi=strategy['fixed_phi'], fixed_nu=strategy['fixed_nu'],
                                     fixed_brlen=strategy['fixed_brlen'])
        score = None if nllh is None else -nllh
        return score, status

    # pickling ships the host state only (no device handles), as solvers
    # cross process boundaries in the reference's parallel search
    def __getstate__(self):
        st = dict(self.__dict__)
        st['_forest'] = None
        st['_forest_key'] = None
        st.pop('_last_estep', None)
        return st


class _DeviceObjective:
    """SLSQP objective over x = [branch lengths (post-order over the trees),
    nu, phi]: the device NLL and its gradient (one inside pass and one
    outside pass per new x; the gradient call at the same x reuses them)."""

    def __init__(self, solver, F, fixed_nu, fixed_phi):
        self.F = F
        self.ne = solver.num_edges
        self.lvl = np.concatenate([solver._node_gids(F, ti) for ti in range(len(solver.trees))])
        self.fixed_nu = fixed_nu
        self.fixed_phi = fixed_phi
        self.d_host = np.zeros(F.num_nodes)
        self.d = torch.zeros(F.num_nodes, dtype=torch.float64, device=F.device)
        self.nu = torch.zeros(1, dtype=torch.float64, device=F.device)
        self.phi = torch.zeros(1, dtype=torch.float64, device=F.device)
        self.x = None
        self.evals = 0

    def _eval(self, x):
        x = np.asarray(x, dtype=np.float64)
        if self.x is not None and np.array_equal(x, self.x):
            return
        ne = self.ne
        self.d_host[self.lvl] = x[:ne]
        self.d.copy_(torch.from_numpy(self.d_host))
        self.nu.fill_(float(x[ne]) if self.fixed_nu is None else float(self.fixed_nu))
        self.phi.fill_(float(x[ne + 1]) if self.fixed_phi is None else float(self.fixed_phi))
        nll = self.F.nll(self.d, self.nu, self.phi)
        gd, gnu, gphi = self.F.grad(self.d, self.nu, self.phi)
        g = np.zeros(len(x))
        g[:ne] = gd.cpu().numpy()[self.lvl]
        if self.fixed_nu is None:
            g[ne] = float(gnu.item())
        if self.fixed_phi is None:
            g[ne + 1] = float(gphi.item())
        self.f = float(nll[0].item())
        self.g = g
        self.x = x.copy()
        self.evals += 1

    def fun(self, x):
        self._eval(x)
        return self.f

    def jac(self, x):
        self._eval(x)
        return self.g


class ML_solver(EM_solver):
    """GPU-backed replacement of laml_libs.ML_solver.ML_solver
    (ML_solver.py:18-377).

    The trees stay as given, polytomies included (x and the tree text are
    over their nodes); the device forest holds each polytomy resolved into
    zero-length fixed branches, whose likelihood is the polytomy's.  A leaf
    entry -1 is missing like '?' (ML_solver.az_partition labels it '?',
    ML_solver.py:139-140).  optimize_one keeps the reference's SciPy SLSQP
    call, bounds and linear constraints (ML_solver.py:338-377); its objective
    is the device NLL and its gradient the device's exact gradient, where the
    reference lets SLSQP difference the objective."""
    solver_name = "ML_solver"

    def __init__(self, treeList, data, prior, params={'nu': EPS, 'phi': EPS}):
        self._read_inputs(treeList, data, prior, params)
        self.solver_name = "ML_solver"
        self.has_polytomy = any(len(v.children) > 2 for t in self.trees for v in t.traverse_postorder())
        for t in self.trees:
            for v in t.traverse_postorder():
                v.polytomy_mark = False
                v.mark_fixed = False

    def _leaf_row(self, label):
        return ['?' if x == -1 else x for x in self.charMtrx[label]]

    def _tree_arrays(self, tree, codes_of, free_of, poly_of):
        return resolved_tree_arrays(tree, codes_of, free_of)

    def ultrametric_constr(self):
        """ML_solver.ultrametric_constr (ML_solver.py:95-112): rows of M with
        M x = 0 over x = [branch lengths, nu, phi] (binary trees)."""
        N = len(self.ini_all())
        M = []
        idx = 0
        cons = {}
        for tree in self.trees:
            for node in tree.traverse_postorder():
                if node.is_leaf():
                    c = [0.] * N
                else:
                    c1, c2 = node.children
                    M.append([x - y for (x, y) in zip(cons[id(c1)], cons[id(c2)])])
                    c = cons[id(c1)]
                c[idx] = 1
                cons[id(node)] = c
                idx += 1
        r0 = cons[id(self.trees[0].root)]
        for tree in self.trees[1:]:
            M.append([x - y for (x, y) in zip(r0, cons[id(tree.root)])])
        return M

    # ML_solver.py:207-222
    def bound_nu(self, fixed_nu=None):
        return (EPS, 10) if fixed_nu is None else (fixed_nu - EPS, fixed_nu + EPS)

    def bound_phi(self, fixed_phi=None):
        return (EPS, 0.99) if fixed_phi is None else (fixed_phi - EPS, fixed_phi + EPS)

    def bound_brlen(self):
        return [self.dmin] * self.num_edges, [self.dmax] * self.num_edges

    def get_bound(self, keep_feasible=False, fixed_phi=None, fixed_nu=None):
        from scipy import optimize

        br_lower, br_upper = self.bound_brlen()
        phi_lower, phi_upper = self.bound_phi(fixed_phi=fixed_phi)
        nu_lower, nu_upper = self.bound_nu(fixed_nu=fixed_nu)
        return optimize.Bounds(br_lower + [nu_lower, phi_lower], br_upper + [nu_upper, phi_upper],
                               keep_feasible=keep_feasible)

    def optimize_one(self, randseed, fixed_phi=None, fixed_nu=None, verbose=1, ultra_constr=False):
        """ML_solver.optimize_one (ML_solver.py:338-377).  Returns (nllh,
        "optimal") or (None, SLSQP's message)."""
        from scipy import optimize
        from scipy.sparse import csr_matrix

        _random.seed(a=randseed)
        x0 = self.ini_all(fixed_phi=fixed_phi, fixed_nu=fixed_nu)
        self.az_partition()
        bounds = self.get_bound(fixed_phi=fixed_phi, fixed_nu=fixed_nu)
        constraints = []
        A, b = [], []
        idx = 0
        for tree in self.trees:
            for node in tree.traverse_postorder():
                if node.mark_fixed:
                    a = [0] * len(x0)
                    a[idx] = 1
                    A.append(a)
                    b.append(node.edge_length)
                idx += 1
        if A:
            constraints.append(optimize.LinearConstraint(csr_matrix(A), b, b, keep_feasible=False))
        if ultra_constr:
            M = self.ultrametric_constr()
            constraints.append(optimize.LinearConstraint(csr_matrix(M), [0] * len(M), [0] * len(M),
                                                         keep_feasible=False))
        F = self._forest_for()
        obj = _DeviceObjective(self, F, fixed_nu, fixed_phi)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            out = optimize.minimize(obj.fun, x0, jac=obj.jac, method="SLSQP",
                                    options={'disp': verbose > 0, 'maxiter': 1000}, bounds=bounds,
                                    constraints=constraints)
        self.slsqp_evals = obj.evals
        self.slsqp_iterations = int(out.nit)
        if out.success:
            self.x2params(out.x, fixed_phi=fixed_phi, fixed_nu=fixed_nu)
            f = float(out.fun)
        else:
            if obj.x is not None:
                self.x2params(obj.x, fixed_phi=fixed_phi, fixed_nu=fixed_nu)
            f = None
        status = "optimal" if out.success else out.message
        return f, status


class fastEM_solver(EM_solver):
    """laml_libs.fastEM_solver.fastEM_solver (fastEM_solver.py:149-364).

    The constructor takes the reference's extra inputs (data
    'ordered_leaf_labels' and 'charMtrx_recode', prior 'Q_recode', as
    runner.py:233-243 prepares them) or computes them with ``parse_tree`` /
    ``parse_data``; ``score_tree`` runs the fastEM model through
    ``laml_amd.fastem.EMOptimizer`` (exact FP64 EM on the device) with the
    reference's start values and always-optimised nu and phi, and the exact
    reference EM for 'final_optimization' (fastEM_solver.py:290-292).

    Options (keywords, not in the reference's constructor):
    m_step -- "exact" (default) or "surrogate": the reference's joint L-BFGS
    surrogate M-step (fast_em/laml.py:62-110), see EMOptimizer;
    local_mask -- which branches a local-scoring strategy (``fixed_brlen``
    with ``nodes_to_recompute``) re-optimises: "named" (default) the branches
    its anchor pairs name; "reference" none, as the reference's comparison
    of node labels with node objects (fastEM_solver.py:307-316) leaves every
    branch fixed."""
    solver_name = "fastEM_solver"

    def __init__(self, treeList, data, prior, params={'nu': 0, 'phi': 0, 'sigma': 0}, *, m_step="exact",
                 local_mask="named"):
        from .fastem import EMOptimizer, parse_data, parse_tree

        if local_mask not in ("named", "reference"):
            raise ValueError("local_mask must be 'named' or 'reference'")
        super().__init__(treeList, data, prior, params={'nu': 0, 'phi': 0, 'sigma': 0})
        self.solver_name = "fastEM_solver"
        self.m_step = m_step
        self.local_mask = local_mask
        self.data = data
        self.prior = prior
        self.myEMOptimizers = []
        self.ordered_leaf_labels = []
        have = 'ordered_leaf_labels' in data and 'charMtrx_recode' in data and 'Q_recode' in prior
        for tidx, tree in enumerate(self.trees):
            if have:
                labels = data['ordered_leaf_labels'][tidx]
                C, Qr = data['charMtrx_recode'][tidx], prior['Q_recode'][tidx]
            else:
                parsed = parse_tree(tree)
                phy = parse_data(parsed, {'charMtrx': self.charMtrx}, {'Q': self.Q})
                labels = parsed['relabeling_vector'][:parsed['num_leaves']]
                C, Qr = phy.character_matrix, phy.mutation_priors
            self.ordered_leaf_labels.append(labels)
            parse_tree(tree, has_branch_mask=False, ordered_leaf_labels=labels)
            self.myEMOptimizers.append(EMOptimizer(Qr, C, m_step=m_step))

    def score_tree(self, strategy={'ultra_constr': False, 'fixed_phi': None, 'fixed_nu': None, 'fixed_brlen': None,
                                   'nodes_to_recompute': None}):
        """fastEM_solver.score_tree (fastEM_solver.py:274-353).  Branches named
        by ``nodes_to_recompute`` (anchor pairs) with positive length are
        re-optimised, all others stay fixed when ``fixed_brlen`` is given
        (the reference compares node labels to node objects there, so it
        re-optimises none; this re-optimises the ones it names unless the
        solver was built with local_mask="reference")."""
        from .fastem import parse_tree

        if 'final_optimization' in strategy:
            return EM_solver.score_tree(self, strategy)
        fixed_phi, fixed_nu = strategy['fixed_phi'], strategy['fixed_nu']
        fixed_brlen = strategy['fixed_brlen']
        nodes_to_recompute = strategy.get('nodes_to_recompute')
        has_branch_mask = fixed_brlen is not None
        results = []
        for tidx, tree in enumerate(self.trees):
            recompute = set()
            if nodes_to_recompute and self.local_mask == "named":
                recompute = {id(v) for v in find_LCAs(tree, list(nodes_to_recompute[tidx]))}
            for node in tree.traverse_postorder():
                node.mark_fixed = True
                if fixed_brlen is None:
                    continue
                if id(node) in recompute and (node.edge_length or 0.0) > 0:
                    node.mark_fixed = False
            parsed = parse_tree(tree, has_branch_mask=has_branch_mask,
                                ordered_leaf_labels=self.ordered_leaf_labels[tidx])
            nu = fixed_nu if fixed_nu else 0.5
            phi = fixed_phi if fixed_phi else 0.5
            # the reference optimises nu and phi here whatever the strategy
            # says (fastEM_solver.py:323-324)
            out = self.myEMOptimizers[tidx].optimize(parsed, parsed['branch_lengths'], nu, phi,
                                                     parsed['branch_mask'], False, False)
            self.params.phi = out['phi']
            self.params.nu = out['nu']
            l2n = {v.label: v for v in tree.traverse_postorder()}
            for i, label in enumerate(parsed['relabeling_vector']):
                l2n[label].edge_length = float(out['branch_lengths'][i])
            results = [-out['nllh']]
        self._forest = None
        return max(results), "optimal"
