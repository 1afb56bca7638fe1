"""A batch split into forests stepped on concurrent HIP streams
(laml_amd.streams.PipelinedBatch): every problem ends with the NLL and
parameters of a forest of its part run on its own, bit for bit, and within
1e-9 of the same problem alone."""
import numpy as np
import pytest
import torch

from laml_amd import sim
from laml_amd.forest import Forest
from laml_amd.streams import PipelinedBatch

pytestmark = pytest.mark.gpu


def test_pipelined_batch_matches_parts_alone():
    probs = [sim.make_problem(300 + 120 * i, 90 - 7 * i, 3, 20, 1.0, 0.1, 0.15, 1800 + i) for i in range(5)]
    specs = [sim.problem_spec(t, q, c) for (t, q, c, _) in probs]
    lengths = [[np.full(t.n_nodes, 0.05)] for (t, _, _, _) in probs]
    pb = PipelinedBatch(specs, parts=2, device=torch.cuda.current_device())
    assert len(pb.forests) == 2 and sorted(i for g in pb.groups for i in g) == list(range(5))
    prm = pb.params(lengths, [0.5] * 5, [0.5] * 5)
    res = pb.em(prm, 30, 1e-8)
    pb.synchronize()
    got = [(r[0].cpu().numpy().copy(), r[1].cpu().numpy().copy(), p[0].cpu().numpy().copy(),
            p[1].cpu().numpy().copy(), p[2].cpu().numpy().copy()) for r, p in zip(res, prm)]
    for j, g in enumerate(pb.groups):
        # the part's forest alone, on the default stream
        F = Forest([specs[i] for i in g], device=torch.cuda.current_device())
        d = torch.tensor(F.lengths_to_level([x for i in g for x in lengths[i]]), dtype=torch.float64, device=F.device)
        nu = torch.full((len(g),), 0.5, dtype=torch.float64, device=F.device)
        phi = torch.full((len(g),), 0.5, dtype=torch.float64, device=F.device)
        nll, it = F.em(d, nu, phi, 30, 1e-8)
        for a, b in zip(got[j], (nll.cpu().numpy(), it.cpu().numpy(), d.cpu().numpy(), nu.cpu().numpy(),
                                 phi.cpu().numpy())):
            np.testing.assert_array_equal(a, b)
        F.close()
        # and each problem alone
        for li, i in enumerate(g):
            F1 = Forest([specs[i]], device=torch.cuda.current_device())
            d1 = torch.tensor(F1.lengths_to_level(lengths[i]), dtype=torch.float64, device=F1.device)
            n1 = torch.full((1,), 0.5, dtype=torch.float64, device=F1.device)
            p1 = torch.full((1,), 0.5, dtype=torch.float64, device=F1.device)
            nll1, _ = F1.em(d1, n1, p1, 30, 1e-8)
            np.testing.assert_allclose(got[j][0][li], nll1.cpu().numpy()[0], rtol=1e-9)
            F1.close()
    pb.close()
