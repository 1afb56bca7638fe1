"""M-step on crafted expected counts: the nu solve at its edges and the
M-step status.

laml_mstep is fed per-branch count rows (S0..S4, raw, as laml_estep writes
them) chosen so that the optimum nu* of EM_solver.Mstep's nu subproblem
(EM_solver.py:392-407) lies at 0+ (no silencing mass: only the eps_s clamp of
EM_solver.py:339-348 keeps it positive), near 1e-6 (where the reference's
case_tests/test1/example3.log run ends), exactly at the bound 1, and where
h'(nu) is flat over most of [0, 1] (every branch at dmax).  The device's
branch lengths and nu are compared with the oracle's exact solutions
(oracle.em_oracle.solve_branch / solve_nu, bisection to full precision):
lengths to 1e-9 relative, nu to 1e-6 relative.  Both device nu solvers run:
the multi-launch evaluation (fewer than 8 problems) and the one-block solve
(8 or more).  A row with an infinite count makes the nu objective NaN; that
problem's M-step must report ",failed_nu" (EM_solver.py:450-459) while the
other problems of the same forest report "optimal".
"""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import DMAX, DMIN, Forest
from laml_amd.solvers import mstep_status
from oracle.em_oracle import solve_branch, solve_nu

pytestmark = pytest.mark.gpu

EPS_NU = 1e-5
EPS_S = 1e-6


def _rows_for(case, n_nodes, K, rng):
    """Raw per-branch counts [n_nodes, 5] of one crafted case."""
    c = np.zeros((n_nodes, 5))
    u = rng.uniform(0.2, 1.0, size=(n_nodes, 5)) * K / 5
    if case == "nu_zero":        # no silencing mass at all
        c[:, [0, 1, 3]] = u[:, [0, 1, 3]]
    elif case == "nu_1e-6":      # silencing mass ~1e-6 of the rest
        c[:, [0, 1, 3]] = u[:, [0, 1, 3]]
        c[:, 2] = u[:, 2] * 1e-6
        c[:, 4] = u[:, 4] * 1e-6
    elif case == "nu_one":       # only silencing mass: nu* at the bound 1
        c[:, [2, 4]] = u[:, [2, 4]]
    elif case == "flat":         # S0 ~ 0: every branch at dmax, h' flat for nu >~ 0.3
        c[:, 1] = u[:, 1]
        c[:, 3] = u[:, 3]
        c[:, 2] = u[:, 2] * 0.05
        c[:, 4] = u[:, 4] * 0.05
    elif case == "generic":
        c[:] = u
    else:
        raise ValueError(case)
    return c


def _oracle(rows_raw, K, nu0):
    rows = []
    for s in rows_raw:
        s = [max(EPS_S, float(x)) for x in s]
        tot = sum(s)
        rows.append([x / tot * K for x in s])
    use2 = sum(s[2] for s in rows) > 0 and nu0 > EPS_NU
    use4 = sum(s[4] for s in rows) > 0 and nu0 > EPS_NU
    d = [solve_branch(s, nu0, use2, use4, DMIN, DMAX) for s in rows]
    return d, solve_nu(rows, d)


def _forest(nprob, n=24, K=30, seed=5):
    specs = []
    for p in range(nprob):
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 20, 1.0, 0.1, 0.1, seed + p)
        specs.append(sim.problem_spec(tree, Q, codes))
    return Forest(specs, device=0)


CASES = ["nu_zero", "nu_1e-6", "nu_one", "flat", "generic"]


@pytest.mark.parametrize("nprob", [1, 3, 8])
@pytest.mark.parametrize("nu0", [0.5, 1e-6, 0.999])
def test_nu_edges_match_oracle(nprob, nu0):
    K = 30
    F = _forest(nprob, K=K)
    rng = np.random.default_rng(17)
    node_prob = F.tree_prob[F.tree_of]
    dev = F.device
    # every case meets every solver path: rotate the cases over the problems
    rounds = len(CASES) if nprob < len(CASES) else 1
    for r in range(rounds):
        case_of = [CASES[(p + r) % len(CASES)] for p in range(nprob)]
        counts = np.zeros((F.num_nodes, 5))
        for p in range(nprob):
            sel = np.nonzero(node_prob == p)[0]
            counts[sel] = _rows_for(case_of[p], len(sel), K, rng)
        d = torch.full((F.num_nodes,), 0.05, dtype=torch.float64, device=dev)
        nu = torch.full((nprob,), nu0, dtype=torch.float64, device=dev)
        phi = torch.full((nprob,), 0.1, dtype=torch.float64, device=dev)
        ct = torch.tensor(counts.reshape(-1), dtype=torch.float64, device=dev)
        rt = torch.full((nprob,), 3.0, dtype=torch.float64, device=dev)
        print("nu edges", nprob, nu0, case_of, flush=True)
        F.mstep(d, nu, phi, counts=ct, rtilde=rt)
        torch.cuda.synchronize()
        st = F.mstep_status()
        dh, nuh = d.cpu().numpy(), nu.cpu().numpy()
        for p in range(nprob):
            sel = np.nonzero(node_prob == p)[0]
            wd, wnu = _oracle(counts[sel], K, nu0)
            np.testing.assert_allclose(dh[sel], wd, rtol=1e-9, err_msg=case_of[p])
            assert math.isclose(nuh[p], wnu, rel_tol=1e-6), (case_of[p], nuh[p], wnu)
            if case_of[p] == "nu_one":
                assert wnu == 1.0 and nuh[p] == 1.0
            elif case_of[p] == "nu_1e-6":
                assert 1e-7 < wnu < 1e-5, wnu
            elif case_of[p] == "nu_zero":
                assert wnu < 1e-5, wnu
            assert st[p] == 0 and mstep_status(st[p]) == (True, "optimal")
    F.close()


@pytest.mark.parametrize("nprob", [2, 8])
def test_failed_nu_status(nprob):
    K = 30
    F = _forest(nprob, K=K, seed=11)
    rng = np.random.default_rng(3)
    node_prob = F.tree_prob[F.tree_of]
    counts = np.zeros((F.num_nodes, 5))
    for p in range(nprob):
        sel = np.nonzero(node_prob == p)[0]
        counts[sel] = _rows_for("generic", len(sel), K, rng)
    bad = nprob - 1
    first_bad = np.nonzero(node_prob == bad)[0][0]
    counts[first_bad, 2] = np.inf  # its renormalised S2 is NaN, and so is h'(nu)
    dev = F.device
    d = torch.full((F.num_nodes,), 0.05, dtype=torch.float64, device=dev)
    nu = torch.full((nprob,), 0.3, dtype=torch.float64, device=dev)
    phi = torch.full((nprob,), 0.1, dtype=torch.float64, device=dev)
    ct = torch.tensor(counts.reshape(-1), dtype=torch.float64, device=dev)
    rt = torch.full((nprob,), 3.0, dtype=torch.float64, device=dev)
    print("failed nu status", nprob, flush=True)
    F.mstep(d, nu, phi, counts=ct, rtilde=rt)
    torch.cuda.synchronize()
    st = F.mstep_status()
    for p in range(nprob):
        if p == bad:
            assert mstep_status(st[p]) == (False, ",failed_nu"), st
        else:
            assert st[p] == 0, st
    # the status is that of the last M-step: a clean step clears it
    counts[first_bad, 2] = 1.0
    ct = torch.tensor(counts.reshape(-1), dtype=torch.float64, device=dev)
    nu.fill_(0.3)
    F.mstep(d, nu, phi, counts=ct, rtilde=rt)
    torch.cuda.synchronize()
    assert (F.mstep_status() == 0).all()
    F.close()
