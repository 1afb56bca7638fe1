"""Forest creation on the device: leaf code rows uploaded packed and padded
by a kernel, site labels computed bottom up.  The rows read back (codes as
labels: an edit state, 0, or -1 for a missing entry) equal the host codes
exactly, internal labels equal the rule recomputed on the host from the
children (a shared edit state, 0 if the children disagree or one is
unedited, the other child's label if one is missing), at segment widths 8,
16, 32 and 64 with ragged site counts; a forest built again in reused
device memory gives a bit-identical likelihood."""
import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest

pytestmark = pytest.mark.gpu


def _label(code):
    return np.where(code > 0, code, np.where(code == 0, 0, -1))


def _combine(a, b):
    out = np.where(a == b, a, 0)
    out = np.where(a < 0, b, out)
    out = np.where(b < 0, a, out)
    return np.where((a == 0) | (b == 0), 0, out)


def _expected_labels(F):
    K = F.tree_K[F.tree_of]
    off = np.concatenate([[0], np.cumsum(K[:F.num_leaves])])
    rows = [_label(F.codes[off[v]:off[v + 1]].astype(np.int64)) for v in range(F.num_leaves)]
    for v in range(F.num_leaves, F.num_nodes):
        rows.append(_combine(rows[F.child0[v]], rows[F.child1[v]]))
    return np.concatenate(rows)


@pytest.mark.parametrize("Ks", [(5, 7), (9, 13, 16), (17, 30), (5, 37, 130)])
def test_rows_and_labels_after_create(Ks):
    specs = []
    for i, K in enumerate(Ks):
        tree, Q, codes, _ = sim.make_problem(40 + 17 * i, K, 4, 20, 1.0, 0.1, 0.2, 311 + i)
        specs.append(sim.problem_spec(tree, Q, codes))
    F = Forest(specs, device=torch.cuda.current_device())
    print("forest create: site counts", Ks, flush=True)
    leaf, internal = F.labels()
    np.testing.assert_array_equal(np.concatenate([leaf, internal]), _expected_labels(F))
    F.close()


def test_rebuilt_forest_bit_identical():
    tree, Q, codes, _ = sim.make_problem(3000, 203, 5, 30, 1.0, 0.1, 0.15, 99)
    spec = sim.problem_spec(tree, Q, codes)
    out = []
    for _ in range(3):
        F = Forest([spec], device=torch.cuda.current_device())
        d = torch.tensor(F.lengths_to_level([tree.length]), dtype=torch.float64, device=F.device)
        nu = torch.full((1,), 0.1, dtype=torch.float64, device=F.device)
        phi = torch.full((1,), 0.15, dtype=torch.float64, device=F.device)
        counts, rt, nll = F.estep(d, nu, phi)
        torch.cuda.synchronize()
        out.append((nll.item(), counts.cpu().numpy().copy()))
        F.close()
    for nll, counts in out[1:]:
        assert nll == out[0][0]
        np.testing.assert_array_equal(counts, out[0][1])
