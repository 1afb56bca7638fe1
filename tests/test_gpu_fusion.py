"""Cherry fusion of the level passes on the GPU: with the height-1 level (the
nodes whose children are both leaves) formed and consumed by its parents, the
likelihood is bit-identical to the unfused passes and the counts agree up to
the order of the sums over sites; across launch shapes, partial site chunks,
batches of problems, the EM loop with its graph, and the adjoint gradient.
Fusion runs with the row kernels (more than 32 sites) and stays off with the
segment kernels.  Entry points that read every node's rows refuse a fused
E-step."""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest
from oracle.em_oracle import OracleEM

pytestmark = pytest.mark.gpu


def _pair(problems):
    """The same forest twice: cherry fusion on and off (the default)."""
    specs = [sim.problem_spec(t, q, c) for (t, q, c, _) in problems]
    dev = torch.cuda.current_device()
    Fa = Forest(specs, device=dev)
    Fb = Forest(specs, device=dev)
    # fusion runs in the level pass mode
    Fa.set_pass_mode(0)
    Fb.set_pass_mode(0)
    assert not Fb.fusion()
    Fa.set_fusion(True)
    rows = max(len(q) for (_, q, _, _) in problems) > 32
    assert Fa.fusion() == rows and not Fb.fusion()
    return Fa, Fb


def _params(F, problems, nu=0.2, phi=0.1, start=None):
    L = [t.length if start is None else np.full(t.n_nodes, start) for (t, _, _, _) in problems]
    d = torch.tensor(F.lengths_to_level(L), dtype=torch.float64, device=F.device)
    P = len(problems)
    return (d, torch.full((P,), nu, dtype=torch.float64, device=F.device),
            torch.full((P,), phi, dtype=torch.float64, device=F.device))


def _estep(F, prm):
    counts, rt, nll = F.estep(*prm)
    return nll.cpu().numpy().copy(), counts.cpu().numpy().copy(), rt.cpu().numpy().copy()


CASES = [  # (leaves, sites, problems): row kernels, partial chunks, segment kernels (no fusion)
    (300, 150, 1), (500, 200, 3), (257, 65, 2), (200, 20, 2), (150, 8, 1)]


@pytest.mark.parametrize("n,K,P", CASES)
def test_fused_estep_equals_unfused_and_oracle(n, K, P):
    probs = [sim.make_problem(n + 37 * i, K, 3, 20, 1.0, 0.1, 0.15, 900 + i) for i in range(P)]
    print("case", n, K, P, flush=True)
    Fa, Fb = _pair(probs)
    prm_a, prm_b = _params(Fa, probs), _params(Fb, probs)
    na, ca, ra = _estep(Fa, prm_a)
    nb, cb, rb = _estep(Fb, prm_b)
    # the inside pass does the same operations: likelihoods are identical
    np.testing.assert_array_equal(na, nb)
    np.testing.assert_allclose(ca, cb, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(ra, rb, rtol=1e-12, atol=1e-13)
    t, q, c, _ = probs[0]
    o = OracleEM([t.newick()], sim.codes_to_charmtrx(t, c), q, nu=0.2, phi=0.1)
    assert math.isclose(float(na[0]), o.negative_llh(), rel_tol=1e-9)
    Fa.close()
    Fb.close()


def test_fused_with_shapes():
    probs = [sim.make_problem(600, 200, 5, 50, 1.0, 0.1, 0.1, 77)]
    Fa, Fb = _pair(probs)
    prm_a, prm_b = _params(Fa, probs), _params(Fb, probs)
    ref = _estep(Fb, prm_b)
    H = Fa.num_levels
    a0, b0, top0 = Fa.get_shapes()
    for cpw, top in [(1, H), (4, H), (2, 2), (1, 3), (3, 4)]:
        Fa.set_shapes(np.full(H, cpw), np.full(H, cpw), top)
        print("shape", cpw, top, flush=True)
        nll, counts, rt = _estep(Fa, prm_a)
        np.testing.assert_array_equal(nll, ref[0])
        np.testing.assert_allclose(counts, ref[1], rtol=1e-11, atol=1e-12)
        np.testing.assert_allclose(rt, ref[2], rtol=1e-12, atol=1e-13)
    Fa.close()
    Fb.close()


def test_fused_em_and_graph_match_unfused():
    probs = [sim.make_problem(400 + 50 * i, 120, 3, 20, 1.0, 0.1, 0.1, 950 + i) for i in range(3)]
    Fa, Fb = _pair(probs)
    out = []
    for F in (Fa, Fb):
        d, nu, phi = _params(F, probs, nu=0.5, phi=0.5, start=0.05)
        nll, it = F.em(d, nu, phi, 25, 0.0)
        out.append((nll.cpu().numpy().copy(), F.level_to_tree_lengths(d.cpu().numpy()), nu.cpu().numpy().copy(),
                    phi.cpu().numpy().copy()))
    np.testing.assert_allclose(out[0][0], out[1][0], rtol=1e-12)
    for la, lb in zip(out[0][1], out[1][1]):
        np.testing.assert_allclose(la, lb, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out[0][2], out[1][2], rtol=1e-9)
    np.testing.assert_allclose(out[0][3], out[1][3], rtol=1e-9)
    # replaying the captured step again from the same start gives the same bits
    d, nu, phi = _params(Fa, probs, nu=0.5, phi=0.5, start=0.05)
    nll2, _ = Fa.em(d, nu, phi, 25, 0.0)
    np.testing.assert_array_equal(nll2.cpu().numpy(), out[0][0])
    Fa.close()
    Fb.close()


def test_fused_gradient_matches_unfused():
    probs = [sim.make_problem(350, 100, 3, 20, 1.0, 0.1, 0.1, 970)]
    Fa, Fb = _pair(probs)
    g = []
    for F in (Fa, Fb):
        d, nu, phi = _params(F, probs)
        F.nll(d, nu, phi)
        gd, gnu, gphi = F.grad(d, nu, phi)
        g.append((F.level_to_tree_lengths(gd.cpu().numpy())[0], gnu.item(), gphi.item()))
    np.testing.assert_allclose(g[0][0], g[1][0], rtol=1e-10, atol=1e-10)
    assert math.isclose(g[0][1], g[1][1], rel_tol=1e-10)
    assert math.isclose(g[0][2], g[1][2], rel_tol=1e-10)
    Fa.close()
    Fb.close()


def test_fusion_eligibility_and_refusals():
    probs = [sim.make_problem(120, 70, 3, 10, 1.0, 0.1, 0.1, 990)]
    Fa, Fb = _pair(probs)
    prm = _params(Fa, probs)
    Fa.estep(*prm)
    with pytest.raises(Exception):
        Fa.posteriors(prm[2])
    # with fusion off the same forest serves the posteriors
    Fa.set_fusion(False)
    Fa.estep(*prm)
    Fa.posteriors(prm[2])
    Fa.set_fusion(True)
    assert Fa.fusion()
    Fa.set_checkpoint(2)
    assert not Fa.fusion()
    Fa.set_checkpoint(0)
    Fa.set_pass_mode(1)
    assert not Fa.fusion()
    Fa.set_pass_mode(0)
    assert Fa.fusion()
    Fa.close()
    Fb.close()
    # a two-leaf tree has its root on the cherry level: no fusion
    two = sim.problem_spec(*sim.make_problem(2, 30, 3, 10, 1.0, 0.1, 0.1, 991)[:3])
    F2 = Forest([two], device=torch.cuda.current_device())
    F2.set_pass_mode(0)
    F2.set_fusion(True)
    assert not F2.fusion()
    F2.close()
