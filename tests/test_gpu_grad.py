"""Gradient of the NLL (hand-written adjoint) on the GPU.

Checked three ways: against the Fisher identity evaluated with the oracle's
expected transition counts (the reference's EM E-step quantities), against
central finite differences (h = 1e-6) of the device NLL, and checkpointed
against full storage of the inside vectors.  Tolerances are in each test.
"""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.autograd import neg_log_likelihood
from laml_amd.forest import Forest
from oracle.em_oracle import OracleEM, postorder

pytestmark = pytest.mark.gpu


def _setup(n, K, seed, nu=0.1, phi=0.1, jitter=0.0):
    tree, Q, codes, _ = sim.make_problem(n, K, 5, 40, 1.0, nu, phi, seed)
    L = tree.length.copy()
    if jitter:
        rng = np.random.default_rng(seed)
        L = L * np.exp(jitter * rng.standard_normal(L.shape))
    spec = sim.problem_spec(tree, Q, codes, lengths=L)
    F = Forest([spec], device=torch.cuda.current_device())
    d = torch.tensor(F.lengths_to_level([L]), dtype=torch.float64, device=F.device)
    nu_t = torch.tensor([nu], dtype=torch.float64, device=F.device)
    phi_t = torch.tensor([phi], dtype=torch.float64, device=F.device)
    return tree, Q, codes, L, F, d, nu_t, phi_t


@pytest.mark.parametrize("n,K,seed", [(250, 30, 3), (400, 50, 9)])
def test_gradient_equals_fisher_identity_on_oracle_counts(n, K, seed):
    tree, Q, codes, L, F, d, nu, phi = _setup(n, K, seed, jitter=0.3)
    F.nll(d, nu, phi)
    gd, gnu, gphi = F.grad(d, nu, phi)
    gd_post = np.empty(F.num_nodes)
    gd_post[F.order] = gd.cpu().numpy()
    T = tree.to_tree()
    for v, x in zip(T.traverse_postorder(), L):
        v.edge_length = float(x)
    o = OracleEM([T.newick()], sim.codes_to_charmtrx(tree, codes), Q, nu=0.1, phi=0.1)
    o.az_partition()
    o.estep()
    x = 0.1
    want_nu = 0.0
    for i, v in enumerate(postorder(o.trees[0])):
        S = [sum(v.a["S"][j]) for j in range(5)]
        dv = v.edge_length
        g = -(1 + x) * S[0] + S[1] * (1 / math.expm1(dv) - x) + (S[2] + S[4]) * x / math.expm1(x * dv) - x * S[3]
        assert abs(-g - gd_post[i]) < 1e-9 * (1 + abs(g)), i
        want_nu += dv * (-S[0] - S[1] - S[3] + (S[2] + S[4]) / math.expm1(x * dv))
    assert math.isclose(-want_nu, gnu.item(), rel_tol=1e-9)
    _, Rt = o.phi_terms()
    nobs = float((codes >= 0).sum())
    want_phi = -(-nobs / (1 - 0.1) + Rt / 0.1)
    print("grad", gnu.item(), -want_nu, gphi.item(), want_phi, flush=True)
    assert math.isclose(want_phi, gphi.item(), rel_tol=1e-9)


def test_gradient_matches_central_differences():
    tree, Q, codes, L, F, d, nu, phi = _setup(300, 24, 17, jitter=0.2)
    F.nll(d, nu, phi)
    gd, gnu, gphi = F.grad(d, nu, phi)
    h = 1e-6

    def nll(dd, nn, pp):
        return F.nll(dd, nn, pp).item()

    rng = np.random.default_rng(0)
    for v in rng.choice(F.num_nodes, 24, replace=False):
        dp = d.clone()
        dm = d.clone()
        dp[v] += h
        dm[v] -= h
        fd = (nll(dp, nu, phi) - nll(dm, nu, phi)) / (2 * h)
        assert abs(fd - gd[v].item()) < 1e-5 * (1 + abs(fd)), (v, fd, gd[v].item())
    fd_nu = (nll(d, nu + h, phi) - nll(d, nu - h, phi)) / (2 * h)
    fd_phi = (nll(d, nu, phi + h) - nll(d, nu, phi - h)) / (2 * h)
    print("fd", fd_nu, gnu.item(), fd_phi, gphi.item(), flush=True)
    assert abs(fd_nu - gnu.item()) < 1e-5 * (1 + abs(fd_nu))
    assert abs(fd_phi - gphi.item()) < 1e-5 * (1 + abs(fd_phi))


def test_autograd_function_on_a_batch():
    specs, Ls = [], []
    for i, (n, K) in enumerate([(60, 12), (150, 40), (90, 7)]):
        tree, Q, codes, _ = sim.make_problem(n, K, 3, 20, 1.0, 0.1, 0.1, 50 + i)
        specs.append(sim.problem_spec(tree, Q, codes))
        Ls.append(tree.length)
    F = Forest(specs, device=torch.cuda.current_device())
    d = torch.tensor(F.lengths_to_level(Ls), dtype=torch.float64, device=F.device, requires_grad=True)
    nu = torch.tensor([0.1, 0.2, 0.05], dtype=torch.float64, device=F.device, requires_grad=True)
    phi = torch.tensor([0.1, 0.3, 0.2], dtype=torch.float64, device=F.device, requires_grad=True)
    w = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device=F.device)
    nll = neg_log_likelihood(F, d, nu, phi)
    F.nll(d.detach() * 1.1, nu.detach(), phi.detach())  # clobber the resident state: backward recomputes
    (nll * w).sum().backward()
    F.nll(d.detach(), nu.detach(), phi.detach())
    gd, gnu, gphi = F.grad(d.detach(), nu.detach(), phi.detach())
    wp = w[F.node_problem()]
    torch.testing.assert_close(d.grad, gd * wp, rtol=0, atol=0)
    torch.testing.assert_close(nu.grad, gnu * w, rtol=0, atol=0)
    torch.testing.assert_close(phi.grad, gphi * w, rtol=0, atol=0)


def test_level_checkpointing_gives_identical_gradients():
    tree, Q, codes, L, F, d, nu, phi = _setup(10000, 16, 20261206)
    full_rows = F.inside_rows()
    F.nll(d, nu, phi)
    g_full = [x.clone() for x in F.grad(d, nu, phi)]
    n_full = F.nll(d, nu, phi).item()
    F.set_checkpoint(16)
    ck_rows = F.inside_rows()
    n_ck = F.nll(d, nu, phi).item()
    g_ck = F.grad(d, nu, phi)
    print("checkpoint rows", full_rows, ck_rows, flush=True)
    assert ck_rows < full_rows
    assert n_ck == n_full
    for a, b in zip(g_full, g_ck):
        torch.testing.assert_close(a, b, rtol=0, atol=0)
    # and the E-step / EM in checkpoint mode matches full storage
    F.set_checkpoint(0)
    dd = d.clone()
    nn = nu.clone()
    pp = phi.clone()
    F.em(dd, nn, pp, 3, 0.0)
    F.set_checkpoint(16)
    d2 = d.clone()
    n2 = nu.clone()
    p2 = phi.clone()
    F.em(d2, n2, p2, 3, 0.0)
    torch.testing.assert_close(dd, d2, rtol=0, atol=0)
    torch.testing.assert_close(nn, n2, rtol=0, atol=0)
