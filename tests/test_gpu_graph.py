"""The EM step replayed from a captured HIP graph equals the same step
launched eagerly, bit for bit, across repeated calls, changed inputs and
changed launch shapes (each replay checked, with eager work in between)."""
import os

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest

pytestmark = pytest.mark.gpu


def _forest(graph, specs):
    old = os.environ.get("LAML_EM_GRAPH")
    os.environ["LAML_EM_GRAPH"] = "1" if graph else "0"
    try:
        return Forest(specs, device=torch.cuda.current_device())
    finally:
        if old is None:
            del os.environ["LAML_EM_GRAPH"]
        else:
            os.environ["LAML_EM_GRAPH"] = old


def _run(F, L, nus, phis, iters, tol):
    d = torch.tensor(F.lengths_to_level(L), dtype=torch.float64, device=F.device)
    nu = torch.tensor(nus, dtype=torch.float64, device=F.device)
    phi = torch.tensor(phis, dtype=torch.float64, device=F.device)
    trace = torch.zeros((iters + 1) * len(nus), dtype=torch.float64, device=F.device)
    nll, it = F.em(d, nu, phi, iters, tol, trace=trace)
    torch.cuda.synchronize()
    return d.cpu().numpy(), nu.cpu().numpy(), phi.cpu().numpy(), nll.cpu().numpy().copy(), it.cpu().numpy().copy(), \
        trace.cpu().numpy()


@pytest.mark.parametrize("K", [20, 150])
def test_graph_replay_equals_eager(K):
    specs, L = [], []
    for i, n in enumerate([300, 120, 800]):
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 40, 1.0, 0.1, 0.1, 500 + i)
        specs.append(sim.problem_spec(tree, Q, codes))
        L.append(np.full(tree.n_nodes, 0.05))
    G = _forest(True, specs)
    E = _forest(False, specs)
    starts = [([0.5, 0.4, 0.3], [0.5, 0.2, 0.1]), ([0.1, 0.1, 0.1], [0.3, 0.3, 0.3])]
    for rep in range(3):
        for nus, phis in starts:
            for iters, tol in [(7, 0.0), (40, 1e-8)]:
                g = _run(G, L, nus, phis, iters, tol)
                torch.zeros(1000, device=G.device).sum().item()  # unrelated eager work between replays
                e = _run(E, L, nus, phis, iters, tol)
                for a, b in zip(g, e):
                    np.testing.assert_array_equal(a, b)
        # a change of launch shapes is picked up (re-captured)
        a0, b0, top = G.get_shapes()
        G.set_shapes(a0, b0, G.num_levels if rep == 0 else top)
        E.set_shapes(a0, b0, E.num_levels if rep == 0 else top)
