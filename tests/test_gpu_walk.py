"""The tree-walk pass mode (one launch per pass, a block per (tree, site
chunk)) against the level-synchronous mode and the oracle: likelihood,
expected counts, gradients and EM trajectories on forests of uneven trees
with every segment width (8 to 64 sites per chunk)."""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest
from oracle.em_oracle import OracleEM

pytestmark = pytest.mark.gpu


def _forest(sizes, K, seed):
    specs, L, o_nll = [], [], []
    for i, n in enumerate(sizes):
        tree, Q, codes, _ = sim.make_problem(n, K, 3, 30, 1.0, 0.1, 0.15, seed + i)
        specs.append(sim.problem_spec(tree, Q, codes))
        L.append(tree.length)
        if n <= 400:
            o = OracleEM([tree.newick()], sim.codes_to_charmtrx(tree, codes), Q, nu=0.1, phi=0.15)
            o_nll.append(o.negative_llh())
        else:
            o_nll.append(None)
    F = Forest(specs, device=torch.cuda.current_device())
    d = torch.tensor(F.lengths_to_level(L), dtype=torch.float64, device=F.device)
    P = len(sizes)
    nu = torch.full((P,), 0.1, dtype=torch.float64, device=F.device)
    phi = torch.full((P,), 0.15, dtype=torch.float64, device=F.device)
    return F, d, nu, phi, o_nll


@pytest.mark.parametrize("K", [6, 13, 30, 64, 130])
def test_walk_equals_levels_and_oracle(K):
    F, d, nu, phi, o_nll = _forest([2, 3, 40, 257, 900, 5, 120], K, 900 + K)
    out = {}
    for mode in (0, 1):
        F.set_pass_mode(mode)
        nll = F.nll(d, nu, phi).cpu().numpy().copy()
        counts, rt, nll2 = F.estep(d, nu, phi)
        gd, gnu, gphi = F.grad(d, nu, phi)
        out[mode] = (nll, nll2.cpu().numpy().copy(), counts.cpu().numpy().copy(), rt.cpu().numpy().copy(),
                     gd.cpu().numpy().copy(), gnu.cpu().numpy().copy(), gphi.cpu().numpy().copy())
    for a, b in zip(out[0], out[1]):
        np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-11)
    for got, want in zip(out[1][0], o_nll):
        if want is not None:
            assert math.isclose(got, want, rel_tol=1e-9)


def test_walk_em_matches_levels():
    F, d, nu, phi, _ = _forest([300, 60, 1500, 9, 700], 90, 77)
    res = {}
    for mode in (0, 1):
        F.set_pass_mode(mode)
        dd, nn, pp = d.clone(), nu.clone(), phi.clone()
        trace = torch.zeros(31 * 5, dtype=torch.float64, device=F.device)
        nll, it = F.em(dd, nn, pp, 30, 1e-8, trace=trace)
        res[mode] = (nll.cpu().numpy().copy(), it.cpu().numpy().copy(), dd.cpu().numpy(), nn.cpu().numpy(),
                     pp.cpu().numpy(), trace.cpu().numpy())
    np.testing.assert_allclose(res[1][0], res[0][0], rtol=1e-10)
    np.testing.assert_array_equal(res[1][1], res[0][1])
    for a, b in zip(res[0][2:], res[1][2:]):
        np.testing.assert_allclose(b, a, rtol=1e-7, atol=1e-9)


def test_walk_rejects_checkpointing():
    F, d, nu, phi, _ = _forest([50, 60], 20, 5)
    F.set_pass_mode(1)
    with pytest.raises(Exception):
        F.set_checkpoint(4)
    F.set_pass_mode(0)
    F.set_checkpoint(4)
    with pytest.raises(Exception):
        F.set_pass_mode(1)


def test_autotuned_plan_gives_the_same_results():
    from laml_amd.autotune import autotune
    F, d, nu, phi, _ = _forest([800, 50, 2000, 300], 140, 31)
    counts0, rt0, nll0 = [x.cpu().numpy().copy() for x in F.estep(d, nu, phi)]
    plan, log = autotune(F, d, nu, phi, reps=2)
    assert len(log) >= 3 and plan.describe()["mode"] in ("levels", "trees", "hybrid")
    counts1, rt1, nll1 = [x.cpu().numpy().copy() for x in F.estep(d, nu, phi)]
    np.testing.assert_allclose(nll1, nll0, rtol=1e-12)
    np.testing.assert_allclose(counts1, counts0, rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(rt1, rt0, rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("K", [13, 64, 130])
def test_hybrid_equals_levels(K):
    """The hybrid mode (level launches below a split level, one tree walk for
    the levels above it, root included) at every split gives the level
    mode's likelihood, counts, gradient and EM trajectory."""
    F, d, nu, phi, o_nll = _forest([2, 3, 40, 257, 900, 5, 120], K, 700 + K)
    H = F.num_levels
    a0, b0, top0 = F.get_shapes()
    F.set_pass_mode(0)
    F.set_shapes(a0, b0, H)
    ref = (F.nll(d, nu, phi).cpu().numpy().copy(),) + tuple(x.cpu().numpy().copy() for x in F.estep(d, nu, phi)) \
        + tuple(x.cpu().numpy().copy() for x in F.grad(d, nu, phi))
    for split in (1, 2, 3, H // 2, H - 1, H):
        F.set_pass_mode(0)
        F.set_shapes(a0, b0, split)
        F.set_pass_mode(2)
        print("hybrid split", split, "of", H, flush=True)
        got = (F.nll(d, nu, phi).cpu().numpy().copy(),) + tuple(x.cpu().numpy().copy() for x in F.estep(d, nu, phi)) \
            + tuple(x.cpu().numpy().copy() for x in F.grad(d, nu, phi))
        for a, b in zip(ref, got):
            np.testing.assert_allclose(b, a, rtol=1e-11, atol=1e-11)
    for got, want in zip(ref[0], o_nll):
        if want is not None:
            assert math.isclose(got, want, rel_tol=1e-9)
    # the EM loop (captured step) in hybrid mode follows the level mode
    res = {}
    for mode, split in ((0, H), (2, max(2, H // 2))):
        F.set_pass_mode(0)
        F.set_shapes(a0, b0, split)
        F.set_pass_mode(mode)
        dd, nn, pp = d.clone(), nu.clone(), phi.clone()
        nll, it = F.em(dd, nn, pp, 15, 0.0)
        res[mode] = (nll.cpu().numpy().copy(), dd.cpu().numpy().copy(), nn.cpu().numpy().copy())
    np.testing.assert_allclose(res[2][0], res[0][0], rtol=1e-11)
    np.testing.assert_allclose(res[2][1], res[0][1], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(res[2][2], res[0][2], rtol=1e-9)
    F.set_pass_mode(0)
    with pytest.raises(Exception):
        F.set_pass_mode(3)
