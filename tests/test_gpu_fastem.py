"""The fastEM model on the device (observed-state alphabets with their raw
prior mass, -1 missing like '?') against the dense per-state oracle, the
EMOptimizer API, and the reference's fastEM known answer on test4
(unit_tests_fastEMSolver.py test_1: |NLL - 1007.009158767873| < 10, phi and nu
0 to 4 places)."""
import math

import numpy as np
import pytest
import torch

from helpers import load_test4
from laml_amd import sim
from laml_amd.fastem import EMOptimizer, fastem_problem, parse_data, parse_tree
from laml_amd.forest import Forest
from laml_amd.solvers import fastEM_solver
from oracle.dense import dense_estep

pytestmark = pytest.mark.gpu


def _fastem_case(n, K, seed, nu=0.15, phi=0.2):
    tree, Q, codes, _ = sim.make_problem(n, K, 3, 15, 1.0, nu, phi, seed)
    T = tree.to_tree()
    msa = sim.codes_to_charmtrx(tree, codes)
    # an explicit -1 in a few entries (missing in the fastEM model)
    rng = np.random.default_rng(seed)
    for c in list(msa)[::7]:
        k = int(rng.integers(K))
        msa[c][k] = -1
    p = parse_tree(T)
    phy = parse_data(p, {"charMtrx": msa}, {"Q": Q})
    return T, p, phy


def _dense_inputs(T, p, phy):
    n = p["num_leaves"]
    msa_r = {p["relabeling_vector"][i]: [int(x) for x in phy.character_matrix[i]] for i in range(n)}
    Qr = [{j + 1: float(phy.mutation_priors[k][j]) for j in range(phy.max_alphabet_size)}
          for k in range(phy.num_characters)]
    return T.newick(), msa_r, Qr


def test_device_estep_on_fastem_model_matches_dense_oracle():
    T, p, phy = _fastem_case(150, 24, 11)
    # the device takes parse_tree's float32 lengths; give the oracle the same
    for v in T.traverse_postorder():
        v.edge_length = float(np.float32(v.edge_length))
    root = int(np.nonzero(p["parent"] < 0)[0][0])
    bl = np.asarray(p["branch_lengths"], dtype=np.float64)
    spec, post = fastem_problem(p["children"], root, bl, p["branch_mask"], phy.character_matrix,
                                phy.mutation_priors, 3)
    assert np.any(spec.sites.qsum < 0.999)  # the model really has unnormalised alphabets
    F = Forest([spec], device=torch.cuda.current_device())
    d = torch.tensor(F.lengths_to_level([spec.trees[0].length]), dtype=torch.float64, device=F.device)
    nu = torch.tensor([0.15], dtype=torch.float64, device=F.device)
    phi = torch.tensor([0.2], dtype=torch.float64, device=F.device)
    counts, rt, nll = F.estep(d, nu, phi)
    nwk, msa_r, Qr = _dense_inputs(T, p, phy)
    # the dense oracle reads float32 priors exactly as the device does
    want_nll, want_counts, want_rt = dense_estep(nwk, msa_r, Qr, 0.15, 0.2, missing=("?", -1), silenced=())
    # device rows by the fastEM tree's post-order (its children lists) ->
    # node labels -> the dense oracle's post-order over the same tree
    got_post = counts.view(-1, 5).cpu().numpy()[F.inv]
    by_label = {p["relabeling_vector"][v]: got_post[i] for i, v in enumerate(post)}
    got = np.asarray([by_label[v.label] for v in T.traverse_postorder()])
    print("fastEM model", nll.item(), want_nll, flush=True)
    assert math.isclose(nll.item(), want_nll, rel_tol=1e-9)
    np.testing.assert_allclose(got, want_counts, rtol=1e-9, atol=1e-10)
    assert math.isclose(rt.item(), want_rt, rel_tol=1e-9)


def test_emoptimizer_api_and_result():
    T, p, phy = _fastem_case(300, 30, 12)
    opt = EMOptimizer(phy.mutation_priors, phy.character_matrix)
    out = opt.optimize(p.get("nxtree", p), p["branch_lengths"], 0.5, 0.5, p["branch_mask"])
    assert set(out) == {"nllh", "branch_lengths", "nu", "phi", "em_iterations"}
    assert out["em_iterations"] >= 2 and 0 < out["nu"] < 1 and 0 <= out["phi"] < 1
    # the returned likelihood is the fastEM model's at the returned parameters
    bl = out["branch_lengths"]
    l2n = {v.label: v for v in T.traverse_postorder()}
    for i, label in enumerate(p["relabeling_vector"]):
        l2n[label].edge_length = float(bl[i])
    nwk, msa_r, Qr = _dense_inputs(T, p, phy)
    want, _, _ = dense_estep(nwk, msa_r, Qr, out["nu"], out["phi"], missing=("?", -1), silenced=())
    assert math.isclose(out["nllh"], want, rel_tol=1e-9)
    # EM improved on the start
    start = dict(zip(p["relabeling_vector"], p["branch_lengths"]))
    for label, v in l2n.items():
        v.edge_length = float(start[label])
    nwk0, _, _ = _dense_inputs(T, p, phy)
    nll0, _, _ = dense_estep(nwk0, msa_r, Qr, 0.5, 0.5, missing=("?", -1), silenced=())
    assert out["nllh"] < nll0


def test_fastem_solver_reference_known_answer_test4():
    T, msa, Q = load_test4()
    s = fastEM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"phi": 0, "nu": 0})
    score, status = s.score_tree(strategy={"ultra_constr": False, "fixed_phi": None, "fixed_nu": None,
                                           "fixed_brlen": None, "nodes_to_recompute": None})
    print("fastEM test4", -score, s.params.phi, s.params.nu, flush=True)
    assert status == "optimal"
    assert abs(-score - 1007.009158767873) < 10
    assert round(s.params.phi - 0, 4) == 0 and round(s.params.nu - 0, 4) == 0


def test_surrogate_em_descends_and_returns_the_reference_values():
    """m_step="surrogate" (fast_em/laml.py:112-194 in FP64).  Its last M-step
    lowers the reference's surrogate from the E-step's point (checked with
    that E-step's counts); the returned likelihood is the fastEM model's at
    the parameters after that M-step, while the returned parameters are
    those before it (the reference's return values); the EM ends below its
    start.  The surrogate's loss and gradient are pinned on the CPU
    (tests/test_fastem_surrogate_host.py); 100 L-BFGS iterations do not
    reach its stationary point here (branches whose optimum is at length
    0 move towards it slowly in log space, as in the reference)."""
    from laml_amd.fastem import surrogate_loss

    T, p, phy = _fastem_case(60, 20, 21)
    opt = EMOptimizer(phy.mutation_priors, phy.character_matrix, m_step="surrogate")
    out = opt.optimize(p.get("nxtree", p), p["branch_lengths"], 0.5, 0.5, p["branch_mask"])
    print("surrogate EM", out["nllh"], out["em_iterations"], out["nu"], out["phi"], flush=True)
    assert out["em_iterations"] >= 2 and 0 < out["nu"] < 1 and 0 < out["phi"] < 1
    root = int(np.nonzero(p["parent"] < 0)[0][0])
    spec, post = fastem_problem(p["children"], root, np.asarray(p["branch_lengths"], dtype=np.float64),
                                p["branch_mask"], phy.character_matrix, phy.mutation_priors, 3)
    F = Forest([spec], device=torch.cuda.current_device())
    x, xn = opt.last_x, opt.last_x_new
    N = spec.trees[0].n_nodes

    def dev(xv):
        d = torch.tensor(F.lengths_to_level([np.exp(xv[:N])]), dtype=torch.float64, device=F.device)
        m = 1.0 / (1.0 + np.exp(-xv[N:]))
        return (d, torch.tensor([m[0]], dtype=torch.float64, device=F.device),
                torch.tensor([m[1]], dtype=torch.float64, device=F.device))

    counts, rt, _ = F.estep(*dev(x))
    S = np.empty((N, 5))
    S[F.order] = counts.view(-1, 5).cpu().numpy()
    C = phy.character_matrix
    args = (S, float((C == -1).sum()), float((C != -1).sum()), float(rt.item()))
    l0, _ = surrogate_loss(x, *args)
    l1, _ = surrogate_loss(xn, *args)
    assert l1 < l0
    # the returned likelihood is the one at x_new; the returned parameters are x
    assert abs(F.nll(*dev(xn)).item() - out["nllh"]) <= 1e-12 * abs(out["nllh"])
    bl_x = np.empty(len(p["children"]))
    for i, v in enumerate(post):
        bl_x[v] = np.exp(x[i])
    np.testing.assert_array_equal(out["branch_lengths"], bl_x)
    assert out["nu"] == 1.0 / (1.0 + np.exp(-x[N])) and out["phi"] == 1.0 / (1.0 + np.exp(-x[N + 1]))
    # EM improved on the start
    l2n = {v.label: v for v in T.traverse_postorder()}
    start = dict(zip(p["relabeling_vector"], p["branch_lengths"]))
    for label, v in l2n.items():
        v.edge_length = float(start[label])
    nwk0, msa_r, Qr = _dense_inputs(T, p, phy)
    nll0, _, _ = dense_estep(nwk0, msa_r, Qr, 0.5, 0.5, missing=("?", -1), silenced=())
    assert out["nllh"] < nll0
    F.close()


def test_fastem_solver_surrogate_known_answer_test4():
    """The reference's fastEM KAT (unit_tests_fastEMSolver.py test_1) with
    the reference's surrogate M-step."""
    T, msa, Q = load_test4()
    s = fastEM_solver([T], {"charMtrx": msa}, {"Q": Q}, {"phi": 0, "nu": 0}, m_step="surrogate")
    score, status = s.score_tree(strategy={"ultra_constr": False, "fixed_phi": None, "fixed_nu": None,
                                           "fixed_brlen": None, "nodes_to_recompute": None})
    print("fastEM surrogate test4", -score, s.params.phi, s.params.nu, flush=True)
    assert status == "optimal"
    assert abs(-score - 1007.009158767873) < 10
    assert round(s.params.phi - 0, 4) == 0 and round(s.params.nu - 0, 4) == 0


@pytest.mark.parametrize("mask", ["named", "reference"])
def test_local_scoring_mask(mask):
    """A local-scoring strategy (fixed_brlen + nodes_to_recompute naming the
    five branches around an NNI move): with local_mask="named" exactly those
    branches move; with "reference" (fastEM_solver.py:307-316 compares node
    labels with node objects) none does."""
    from laml_amd.lca import find_LCAs
    from laml_amd.tree import read_tree_newick

    tree, Q, codes, _ = sim.make_problem(40, 15, 3, 10, 1.0, 0.1, 0.1, 77)
    msa = sim.codes_to_charmtrx(tree, codes)
    T = tree.to_tree()
    nwk = T.newick()
    s = fastEM_solver([nwk], {"charMtrx": msa}, {"Q": Q}, {"phi": 0.1, "nu": 0.1}, local_mask=mask)
    t = s.trees[0]
    # anchor pairs of every branch (first leaf of the first and last child)
    for v in t.traverse_postorder():
        v.anchors = (v.label, v.label) if v.is_leaf() else (v.children[0].anchors[0], v.children[-1].anchors[0])
    inner = [v for v in t.traverse_postorder() if not v.is_leaf() and v.parent is not None]
    u = inner[len(inner) // 2]
    named = [u] + list(u.children) + [u.parent] + [c for c in u.parent.children if c is not u]
    fixed = {v.anchors: v.edge_length for v in t.traverse_postorder()}
    strategy = {"ultra_constr": False, "fixed_phi": 0.1, "fixed_nu": 0.1, "fixed_brlen": [fixed],
                "nodes_to_recompute": [[v.anchors for v in named]]}
    before = {v.anchors: v.edge_length for v in t.traverse_postorder()}
    score, status = s.score_tree(strategy=strategy)
    after = {v.anchors: v.edge_length for v in s.trees[0].traverse_postorder()}
    # parse_tree hands the lengths over in float32 (as the reference's jnp
    # arrays), so a fixed branch comes back rounded to float32
    moved = {a for a in before if abs(after[a] - before[a]) > 1e-6 * max(1.0, before[a])}
    print(mask, "moved", len(moved), "score", score, flush=True)
    if mask == "named":
        assert moved and moved <= {v.anchors for v in named}
    else:
        assert not moved
