"""The inside pass's level-1 kernel (nodes whose children are both leaves; a
wavefront's chunks loaded before any store) at every chunk count per
wavefront it specialises (1..8) and past it (the generic row path), on a tree
whose rows hold 11 chunks of 64 sites (700 sites).  The inside pass sums
nothing over sites, so the likelihood and the E-step counts are bit-identical
across the shapes; they equal the C oracle's (NLL 1e-9 relative, counts
1e-9 relative to the branch's largest count)."""
import math

import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest
from oracle import pmm_ref

pytestmark = pytest.mark.gpu


def test_level1_chunks_per_wavefront_bit_identical_and_oracle():
    tree, Q, codes, _ = sim.make_problem(300, 700, 3, 20, 1.0, 0.1, 0.15, 4242)
    F = Forest([sim.problem_spec(tree, Q, codes)], device=torch.cuda.current_device())
    assert -(-int(F.tree_K.max()) // 64) == 11 and F.chains() == 0
    d = torch.tensor(F.lengths_to_level([tree.length]), dtype=torch.float64, device=F.device)
    nu = torch.full((1,), 0.1, dtype=torch.float64, device=F.device)
    phi = torch.full((1,), 0.15, dtype=torch.float64, device=F.device)
    a0, b0, top0 = F.get_shapes()
    results = []
    for c in [1, 2, 3, 4, 5, 6, 7, 8, 11]:
        a = a0.copy()
        a[1] = c
        F.set_shapes(a, b0, top0)
        counts, rt, nll = F.estep(d, nu, phi)
        torch.cuda.synchronize()
        results.append((c, nll.item(), counts.cpu().numpy().copy(), rt.item()))
        print("level-1 chunks per wavefront", c, nll.item(), flush=True)
    for c, nll, counts, rt in results[1:]:
        assert nll == results[0][1], c
        np.testing.assert_array_equal(counts, results[0][2])
        assert rt == results[0][3]
    ref = pmm_ref.RefProblem(tree.c0, tree.c1, tree.parent, codes, Q, None, 1.0 / (len(Q) * codes.shape[0]))
    nll_ref, counts_ref, _, _ = ref.estep(tree.length, 0.1, 0.15)
    assert math.isclose(results[0][1], nll_ref, rel_tol=1e-9)
    S = np.empty((F.num_nodes, 5))
    S[F.order] = results[0][2].reshape(-1, 5)
    scale = np.maximum(np.abs(counts_ref).max(axis=1, keepdims=True), 1e-300)
    assert np.abs(S - counts_ref).max() <= 1e-9 * scale.max()
    assert (np.abs(S - counts_ref) <= 1e-9 * scale).all()
    F.close()
