"""laml_posteriors_noalloc (the posteriors entry point without allocation,
staging in a fixed device array) against laml_posteriors through the C ABI:
bit-identical posteriors on two calls in a row, the second on a smaller
forest (the array keeps the first call's contents), and a clear refusal of a
forest too large for the array."""
import ctypes

import numpy as np
import pytest
import torch

from laml_amd import _lib, sim
from laml_amd.forest import Forest

pytestmark = pytest.mark.gpu

F64P = ctypes.POINTER(ctypes.c_double)


def _estep_forest(n, K, seed, nprob=1):
    specs, lens = [], []
    for i in range(nprob):
        tree, Q, codes, _ = sim.make_problem(n + 37 * i, K, 3, 30, 1.0, 0.1, 0.2, seed + i)
        specs.append(sim.problem_spec(tree, Q, codes))
        lens.append(tree.length)
    F = Forest(specs, device=torch.cuda.current_device())
    F.set_fusion(False)
    d = torch.tensor(F.lengths_to_level(lens), dtype=torch.float64, device=F.device)
    nu = torch.full((nprob,), 0.1, dtype=torch.float64, device=F.device)
    phi = torch.full((nprob,), 0.2, dtype=torch.float64, device=F.device)
    F.estep(d, nu, phi)
    return F, phi


def _noalloc(F, phi):
    K = F.tree_K[F.tree_of]
    need = int(K.sum())
    p0 = np.full(need, np.nan)
    pS = np.full(need, np.nan)
    rc = _lib.load().laml_posteriors_noalloc(F.handle, _lib.ptr(phi), p0.ctypes.data_as(F64P),
                                             pS.ctypes.data_as(F64P), need, F.stream())
    return rc, p0, pS


def _check(F, phi):
    (l0, i0), (lS, iS) = F.posteriors(phi)
    rc, p0, pS = _noalloc(F, phi)
    _lib.check(rc)
    np.testing.assert_array_equal(p0, np.concatenate([l0, i0]))
    np.testing.assert_array_equal(pS, np.concatenate([lS, iS]))
    return p0


def test_noalloc_posteriors_equal_posteriors_twice_then_smaller():
    print("noalloc: 3 problems x ~1500 leaves x 120 sites", flush=True)
    Fa, pa = _estep_forest(1500, 120, 71, nprob=3)
    _check(Fa, pa)
    print("noalloc: second call, 200 leaves x 40 sites", flush=True)
    Fb, pb = _estep_forest(200, 40, 72)
    _check(Fb, pb)
    # and the larger forest again after the smaller one
    _check(Fa, pa)
    Fa.close()
    Fb.close()


def test_noalloc_refuses_forest_beyond_its_array():
    print("noalloc: refusal of 10000 leaves x 300 sites", flush=True)
    F, phi = _estep_forest(10000, 300, 73)
    rc, p0, _ = _noalloc(F, phi)
    assert rc == 1
    assert b"staging area" in _lib.load().laml_last_error()
    assert np.isnan(p0).all()
    F.close()
