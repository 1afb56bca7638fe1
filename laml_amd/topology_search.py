"""NNI topology search driving the device solvers.

``Topology_search`` follows laml_libs/Topology_search.py: annealed acceptance
(:67-80), one NNI round = shuffle the internal non-root branches and try the
two moves of each until one is accepted (:221-240), local branch-length
optimisation that frees only the five branches around the move and fixes nu,
phi and every other branch by leaf-pair anchors (:242-322), polytomy marking
(:53-65), repeated searches and the final full re-optimisation (:108-173),
and the checkpoint text (:210-216).  Scoring goes through the solver class it
is given (``laml_amd.solvers.EM_solver`` runs on the GPU).

``Topology_search_batched`` replaces the reference's process pool
(laml_libs/Topology_search_parallel.py:13-56, batches of 16 candidates over
``multiprocessing.Pool``) with device batches: every NNI candidate of a batch
becomes one problem of a single device forest and all of them run EM
together, each stopping at its own convergence test (``score_batch``).
"""
import random as _random
from copy import deepcopy
from math import exp

import numpy as np
import torch

from ._lib import LAML_OPT_NU, LAML_OPT_PHI
from .forest import Forest
from .lca import find_LCAs
from .solvers import CONV_EPS, DEFAULT_STRATEGY, EM_solver, mstep_status
from .tree import read_tree_newick

CHKPT_FREQ = 10  # laml_libs/core_constants.py:9


def _nni_swap(u, v, u_child, w):
    """Move u_child under v and w under u (Topology_search.py:289-297)."""
    u_child.set_parent(v)
    u.remove_child(u_child)
    v.add_child(u_child)
    w.set_parent(u)
    v.remove_child(w)
    u.add_child(w)


def _nni_unswap(u, v, u_child, w):
    u_child.set_parent(u)
    v.remove_child(u_child)
    u.add_child(u_child)
    w.set_parent(v)
    u.remove_child(w)
    v.add_child(w)


class Topology_search:
    def __init__(self, treeTopoList, solver, data={}, prior={}, params={}, T_cooldown=20, alpha_cooldown=0.9):
        self.treeTopoList = treeTopoList
        self.solver = solver
        self.params = params
        self.data = data
        self.prior = prior
        self.__renew_treeList_obj__()
        self.has_polytomy = any(len(v.children) > 2 for t in self.treeList_obj for v in t.traverse_preorder())
        self.treeTopoList = [t.newick() for t in self.treeList_obj]
        self.T_cooldown = T_cooldown
        self.alpha_cooldown = alpha_cooldown
        # annealing schedule T(t) = a alpha^t + b with T(0) = 1, T(T_cooldown) = 0
        self.b = 1 / (1 - 1 / self.alpha_cooldown ** self.T_cooldown)
        self.a = -self.b / self.alpha_cooldown ** self.T_cooldown

    def __renew_treeList_obj__(self):
        self.treeList_obj = []
        for topo in self.treeTopoList:
            t = read_tree_newick(topo)
            t.suppress_unifurcations()
            self.treeList_obj.append(t)

    def get_solver(self):
        return self.solver(self.treeTopoList, self.data, self.prior, self.params)

    def update_from_solver(self, mySolver):
        self.treeTopoList = mySolver.get_tree_newick()
        self.params = mySolver.get_params()
        if getattr(self.solver, "solver_name", None) == "fastEM_solver":
            # Topology_search.py:46-51: the fastEM path re-reads the trees
            self.__renew_treeList_obj__()

    def __mark_polytomies__(self, eps_len=1e-3):
        """Resolve every polytomy; the created branches get ``mark`` True and
        length eps_len (Topology_search.py:53-65)."""
        self.has_polytomy = False
        for t in self.treeList_obj:
            for v in t.traverse_preorder():
                v.mark = False
                if len(v.children) > 2:
                    self.has_polytomy = True
            t.resolve_polytomies()
        for t in self.treeList_obj:
            for v in t.traverse_preorder():
                if not hasattr(v, "mark"):
                    v.mark = True
                    v.edge_length = eps_len
        self.treeTopoList = [t.newick() for t in self.treeList_obj]

    def __accept_proposal__(self, curr_score, new_score, t):
        if new_score > curr_score:
            return True
        T = max(1e-12, self.a * self.alpha_cooldown ** t + self.b)
        p = min(exp((new_score - curr_score - 1e-12) / T), 1)
        return _random.random() < p

    # ------------------------------------------------------------------
    def search(self, resolve_polytomies=True, maxiter=100, verbose=False, nreps=1, strategy=DEFAULT_STRATEGY,
               checkpoint_file=None):
        original_topos = self.treeTopoList
        original_params = self.params
        best_trees, best_score, best_params = None, -float("inf"), None
        for i in range(nreps):
            if verbose:
                print("Performing nni-search " + str(i + 1), flush=True)
            self.treeTopoList = original_topos
            self.params = original_params
            self.__renew_treeList_obj__()
            if resolve_polytomies:
                self.__mark_polytomies__()
            if strategy['resolve_search_only']:
                if self.has_polytomy:
                    trees, score, params = self.__search_one__(strategy, maxiter=maxiter, verbose=verbose,
                                                               only_marked=True, checkpoint_file=checkpoint_file)
                else:
                    mySolver = self.get_solver()
                    st = deepcopy(strategy)
                    st['fixed_brlen'] = None
                    st['nodes_to_recompute'] = None
                    score, status = mySolver.score_tree(strategy=st)
                    self.update_from_solver(mySolver)
                    trees, params = self.treeTopoList, self.params
            else:
                trees, score, params = self.__search_one__(strategy, maxiter=maxiter, verbose=verbose,
                                                           only_marked=False, checkpoint_file=checkpoint_file)
            # final optimisation of all parameters on the best topology
            self.treeTopoList = trees
            self.params = params
            self.__renew_treeList_obj__()
            mySolver = self.get_solver()
            st = deepcopy(strategy)
            st['fixed_brlen'] = None
            st['final_optimization'] = True
            st['nodes_to_recompute'] = None
            score, status = mySolver.score_tree(strategy=st)
            self.update_from_solver(mySolver)
            trees, params = self.treeTopoList, self.params
            if verbose:
                print("Optimal score for this search: " + str(score), flush=True)
            if score > best_score:
                best_score, best_trees, best_params = score, trees, params
        self.treeTopoList = best_trees
        self.params = best_params
        return best_trees, best_score, best_params

    def __search_one__(self, strategy, maxiter=100, verbose=False, only_marked=False, checkpoint_file=None):
        mySolver = self.get_solver()
        st = deepcopy(strategy)
        st['fixed_brlen'] = None
        st['nodes_to_recompute'] = None
        curr_score, status = mySolver.score_tree(strategy=st)
        self.update_from_solver(mySolver)
        best_score, best_trees, best_params = curr_score, self.treeTopoList, self.params
        for nni_iter in range(maxiter):
            new_score, n_attempts, success = self.single_nni(curr_score, nni_iter, strategy, only_marked=only_marked)
            if not success:
                break
            curr_score = new_score
            if curr_score > best_score:
                best_score, best_trees, best_params = curr_score, self.treeTopoList, self.params
            if verbose:
                print("NNI Iter: %d, checked %d trees, current score %s" % (nni_iter, n_attempts, curr_score),
                      flush=True)
            if nni_iter % CHKPT_FREQ == 0 and checkpoint_file is not None:
                with open(checkpoint_file, "a") as f:
                    f.write(f"NNI Iteration: {nni_iter}\n")
                    f.write(f"Current newick tree: {best_trees}\n")
                    f.write(f"Current negative-llh: {best_score}\n")
                    f.write(f"Current dropout rate: {best_params['phi']}\n")
                    f.write(f"Current silencing rate: {best_params['nu']}\n")
        return best_trees, best_score, best_params

    def _nni_branches(self, only_marked):
        out = []
        for t in self.treeList_obj:
            for v in t.traverse_preorder():
                if v.is_leaf() or v.is_root():
                    continue
                if not only_marked or getattr(v, "mark", False):
                    out.append(v)
        return out

    def single_nni(self, curr_score, nni_iter, strategy, only_marked=False):
        branches = self._nni_branches(only_marked)
        _random.shuffle(branches)
        took = False
        score = -float("inf")
        n_attempts = 0
        while not took and branches:
            u = branches.pop()
            took, score = self.apply_nni(u, curr_score, nni_iter, strategy)
            n_attempts += 2
        return score, n_attempts, took

    def _local_strategy(self, u, strategy):
        """Strategy of one NNI around branch u (Topology_search.py:247-287):
        nu and phi fixed, all branches except u's children, v's children and
        v fixed at their current lengths through leaf-pair anchors."""
        v = u.get_parent()
        st = deepcopy(strategy)
        st['fixed_brlen'] = None
        st['nodes_to_recompute'] = None
        if not strategy['local_brlen_opt']:
            return st
        st['fixed_nu'] = self.params['nu']
        st['fixed_phi'] = self.params['phi']
        free = set(u.child_nodes() + v.child_nodes() + [v])
        anchors_fixed = [[] for _ in self.treeList_obj]
        anchors_free = [[] for _ in self.treeList_obj]
        for ti, t in enumerate(self.treeList_obj):
            for x in t.traverse_postorder():
                if x.is_leaf():
                    x.anchors = (x.label, x.label)
                else:
                    C = x.child_nodes()
                    x.anchors = (C[0].anchors[0], C[-1].anchors[0])
                (anchors_free if x in free else anchors_fixed)[ti].append(x.anchors)
        fixed_brlen = [{} for _ in self.treeList_obj]
        for ti, topo in enumerate(self.treeTopoList):
            cur = read_tree_newick(topo)
            for anc, node in zip(anchors_fixed[ti], find_LCAs(cur, anchors_fixed[ti])):
                fixed_brlen[ti][anc] = node.edge_length
        st['fixed_brlen'] = fixed_brlen
        st['nodes_to_recompute'] = anchors_free
        return st

    def apply_nni(self, u, curr_score, nni_iter, strategy):
        """Try both NNI moves around branch u (Topology_search.py:242-322)."""
        v = u.get_parent()
        w = next(x for x in v.child_nodes() if x is not u)
        u_children = u.child_nodes()
        _random.shuffle(u_children)
        st = self._local_strategy(u, strategy)
        for u_child in u_children:
            _nni_swap(u, v, u_child, w)
            mySolver = self.solver([t.newick() for t in self.treeList_obj], self.data, self.prior, self.params)
            new_score, status = mySolver.score_tree(strategy=st)
            if status != "optimal" and strategy['local_brlen_opt']:
                # as in the reference, the fallback to a full re-optimisation
                # stays in effect for the second move of this branch
                st['fixed_brlen'] = None
                st['nodes_to_recompute'] = None
                mySolver = self.solver([t.newick() for t in self.treeList_obj], self.data, self.prior, self.params)
                new_score, status = mySolver.score_tree(strategy=st)
            if self.__accept_proposal__(curr_score, new_score, nni_iter):
                self.update_from_solver(mySolver)
                return True, new_score
            _nni_unswap(u, v, u_child, w)
        return False, curr_score


# ----------------------------------------------------------------------------
# batched candidate scoring on the device
# ----------------------------------------------------------------------------
def score_batch(jobs, data, prior, params, solver_cls=EM_solver, max_iter=1000, tol=CONV_EPS):
    """Score many candidate trees in one device forest.

    jobs: list of (newick list, strategy) as the reference passes to
    ``solver(...).score_tree(strategy)``.  Each candidate is set up exactly as
    ML_solver.optimize(initials=1, ...) does (ML_solver.py:256-306: one seed
    drawn from the global generator, fixed branches marked through anchors,
    ini_all, x2params) and becomes one problem of the forest; EM runs on all of
    them together with a per-problem convergence test (EM_solver.py:461-512).
    Returns a list of (solver, score, status).  Candidates with the
    ultrametric constraint, and every candidate of a solver other than
    EM_solver (whose score_tree is not this EM), are scored one by one.

    The global random generator is left as it was: the reference scores a
    batch in forked pool workers (Topology_search_parallel.py:25-26), so
    each candidate draws its seed from a copy of the caller's generator and
    the caller's own draws (move shuffles, annealed acceptance) do not
    depend on the scoring.  Each candidate here starts from the caller's
    state, as a candidate alone in a fresh worker does."""
    out = [None] * len(jobs)
    batch = []
    batchable = getattr(solver_cls, "solver_name", None) == "EM_solver"
    caller_state = _random.getstate()
    try:
        for j, (nwks, st) in enumerate(jobs):
            _random.setstate(caller_state)
            _prepare_candidate(j, nwks, st, data, prior, params, solver_cls, batchable, out, batch)
    finally:
        _random.setstate(caller_state)
    _score_prepared(batch, out, max_iter, tol)
    return out


def _prepare_candidate(j, nwks, st, data, prior, params, solver_cls, batchable, out, batch):
    """score_batch's set-up of candidate j: scored at once if it is not
    batchable, else its start drawn and queued in `batch`."""
    s = solver_cls(nwks, data, prior, params)
    if st.get('ultra_constr') or not batchable:
        sc, status = s.score_tree(strategy=st)
        out[j] = (s, sc, status)
        return
    fixed_phi, fixed_nu, fixed_brlen = st['fixed_phi'], st['fixed_nu'], st['fixed_brlen']
    randseed = int(_random.random() * 10000)
    s.em_seed = randseed
    for t, tree in enumerate(s.trees):
        for x in tree.traverse_postorder():
            x.mark_fixed = False
        if fixed_brlen is None:
            continue
        nodes = find_LCAs(tree, list(fixed_brlen[t].keys()))
        for i, (a, b) in enumerate(fixed_brlen[t]):
            nodes[i].edge_length = fixed_brlen[t][(a, b)]
            nodes[i].mark_fixed = True
    _random.seed(a=randseed)
    x0 = s.ini_all(fixed_phi=fixed_phi, fixed_nu=fixed_nu)
    s.x2params(x0, fixed_phi=fixed_phi, fixed_nu=fixed_nu)
    flags = (LAML_OPT_PHI if fixed_phi is None else 0) | (LAML_OPT_NU if fixed_nu is None else 0)
    batch.append((j, s, flags))


def _score_prepared(batch, out, max_iter, tol):
    """EM on the queued candidates as one device forest."""
    if batch:
        specs = [s._problem_spec(flags) for (_, s, flags) in batch]
        dev = torch.cuda.current_device()
        F = Forest(specs, device=dev, dmin=batch[0][1].dmin, dmax=batch[0][1].dmax)
        lengths = [np.asarray([0.0 if v.edge_length is None else float(v.edge_length)
                               for v in t.traverse_postorder()]) for (_, s, _) in batch for t in s.trees]
        d = torch.tensor(F.lengths_to_level(lengths), dtype=torch.float64, device=F.device)
        nu = torch.tensor([float(s.params.nu) for (_, s, _) in batch], dtype=torch.float64, device=F.device)
        phi = torch.tensor([float(s.params.phi) for (_, s, _) in batch], dtype=torch.float64, device=F.device)
        nll, iters = F.em(d, nu, phi, max_iter, tol)
        nll = nll.cpu().numpy()
        mst = F.mstep_status()
        nu_h, phi_h = nu.cpu().numpy(), phi.cpu().numpy()
        per_tree = F.level_to_tree_lengths(d.cpu().numpy())
        ti = 0
        for k, (j, s, _) in enumerate(batch):
            for t in s.trees:
                for v, x in zip(t.traverse_postorder(), per_tree[ti]):
                    if not (v.polytomy_mark or v.mark_fixed):
                        v.edge_length = float(x)
                ti += 1
            s.params.nu = float(nu_h[k])
            s.params.phi = float(phi_h[k])
            processed = []
            for t in s.trees:
                tc = read_tree_newick(t.newick())
                tc.collapse_short_branches(s.dmin * 0.01)
                processed.append(tc.newick())
            s.trees = [read_tree_newick(x) for x in processed]
            for t in s.trees:
                for v in t.traverse_postorder():
                    v.polytomy_mark = False
                    v.mark_fixed = False
            s._forest = None
            out[j] = (s, -float(nll[k]), mstep_status(mst[k])[1])
        F.close()
    return out


class Topology_search_batched(Topology_search):
    """NNI search whose candidates are scored in device batches (the role of
    Topology_search_parallel.py:13-56, which scores batches of 16 candidates
    in a process pool)."""

    def __init__(self, *args, batch_size=16, quartet=True, **kwargs):
        super().__init__(*args, **kwargs)
        self.batch_size = batch_size
        self.quartet = quartet

    def _list_moves(self, only_marked=False):
        """The NNI moves of list_all_nni in the same order (and with the same
        draws from the random generator), without writing tree texts."""
        branches = self._nni_branches(only_marked)
        _random.shuffle(branches)
        moves = []
        for u in branches:
            v = u.get_parent()
            w = next(x for x in v.child_nodes() if x is not u)
            u_children = u.child_nodes()
            _random.shuffle(u_children)
            for u_child in u_children:
                moves.append((u, v, u_child, w))
        return moves

    def _single_nni_quartet(self, curr_score, nni_iter, strategy, only_marked):
        """Every candidate of the round scored by quartet-local EM in one
        launch, then accepted in queue order (the queue order and the
        acceptance rule of Topology_search_parallel.single_nni)."""
        from .quartet import QuartetScorer
        moves = self._list_moves(only_marked)
        if not moves:
            return curr_score, 0, False
        scorer = QuartetScorer(self, strategy, self.solver)
        try:
            results = scorer.score(moves)
            for move, (sc, lazy) in zip(moves, results):
                if self.__accept_proposal__(curr_score, sc, nni_iter):
                    res = lazy.materialise()
                    _nni_swap(*move)
                    self.update_from_solver(res)
                    return sc, len(moves), True
        finally:
            scorer.close()
        return curr_score, len(moves), False

    def list_all_nni(self, strategy, only_marked=False):
        """Every NNI move of the current trees as (newicks, strategy, move)
        (Topology_search_parallel.py:65-130)."""
        branches = self._nni_branches(only_marked)
        _random.shuffle(branches)
        moves = []
        for u in branches:
            v = u.get_parent()
            w = next(x for x in v.child_nodes() if x is not u)
            st = self._local_strategy(u, strategy)
            u_children = u.child_nodes()
            _random.shuffle(u_children)
            for u_child in u_children:
                _nni_swap(u, v, u_child, w)
                moves.append(([t.newick() for t in self.treeList_obj], st, (u, v, u_child, w)))
                _nni_unswap(u, v, u_child, w)
        return moves

    def single_nni(self, curr_score, nni_iter, strategy, only_marked=False):
        """Topology_search_parallel.py:13-54: walk the move queue in batches,
        accept the first candidate that passes the annealed test; candidates
        that did not reach status 'optimal' are queued once more with free
        branch lengths, and the first batch of that queue is tried after the
        main queue is exhausted."""
        if (self.quartet and strategy['local_brlen_opt'] and not strategy['ultra_constr']
                and getattr(self.solver, "solver_name", None) == "EM_solver"):
            return self._single_nni_quartet(curr_score, nni_iter, strategy, only_marked)
        queue = self.list_all_nni(strategy, only_marked=only_marked)
        N = len(queue)
        retry = []
        checked_all = False
        took = False
        new_score = curr_score
        start, end = 0, min(self.batch_size, N)
        while True:
            chunk = queue[start:end]
            results = score_batch([(m[0], m[1]) for m in chunk], self.data, self.prior, self.params,
                                  solver_cls=self.solver) if chunk else []
            for (nwks, st, move), (s, sc, status) in zip(chunk, results):
                if status == "optimal":
                    new_score = sc
                    if self.__accept_proposal__(curr_score, sc, nni_iter):
                        _nni_swap(*move)
                        self.update_from_solver(s)
                        took = True
                        break
                elif not checked_all:
                    st['fixed_brlen'] = None
                    retry.append((nwks, st, move))
            if checked_all or took:
                break
            if end == N:
                checked_all = True
                queue = retry
                start = 0
            else:
                start += self.batch_size
            end = min(start + self.batch_size, len(queue))
        return new_score, end, took
