"""The C/OpenMP restatement of the EM oracle (oracle/pmm_ref.c, the CPU
baseline) against the Python oracle and the reference's known answers: NLL
constants, E-step counts and phi sums, M-step updates, EM iterations, and the
reference's end-to-end EM KAT.  CPU only."""
import math
import random

import numpy as np
import pytest

from helpers import kat_inputs, load_example1, load_kats, load_test4
from oracle import pmm_ref
from oracle.em_oracle import OracleEM, postorder
from test_oracle import _random_tiny_case

KATS = load_kats()


@pytest.fixture(scope="module", autouse=True)
def _built():
    try:
        pmm_ref.lib()
    except RuntimeError as e:  # pragma: no cover
        pytest.fail(str(e))


@pytest.mark.parametrize("case", KATS["nllh_kats"], ids=lambda c: c["name"])
def test_c_oracle_reference_constants(case):
    T, msa, Q, nu, phi = kat_inputs(case)
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    p, d = pmm_ref.RefProblem.from_oracle(o)
    assert abs(p.nll(d, nu, phi) - case["nllh"]) < 2e-9


def _estep_python(o):
    o.az_partition()
    o.estep()
    nodes = [v for r in o.trees for v in postorder(r)]
    counts = np.asarray([[sum(v.a["S"][j]) for j in range(5)] for v in nodes])
    R, Rt = o.phi_terms()
    return -sum(sum(r.a["L0"]) for r in o.trees), counts, R, Rt


@pytest.mark.parametrize("seed", range(12))
def test_c_estep_matches_python_oracle(seed):
    rng = random.Random(1000 + seed)
    n = 4 + seed % 5
    T, msa, Q = _random_tiny_case(rng, n, 6, 3)
    nu = [0.02, 0.05, 0.3, 0.9][seed % 4]
    phi = [0.0, 0.1, 0.35][seed % 3]
    o = OracleEM([T], msa, Q, nu=nu, phi=phi)
    nll, counts, R, Rt = _estep_python(o)
    p, d = pmm_ref.RefProblem.from_oracle(o)
    cn, cc, cR, cRt = p.estep(d, nu, phi)
    assert math.isclose(cn, nll, rel_tol=1e-12, abs_tol=1e-12)
    assert np.allclose(cc, counts, rtol=1e-10, atol=1e-12)
    assert cR == R
    assert math.isclose(cRt, Rt, rel_tol=1e-10, abs_tol=1e-12)


def test_c_estep_and_em_match_python_oracle_on_example1():
    T, msa, Q = load_example1()
    o = OracleEM([T], msa, Q, nu=0.2, phi=0.1)
    nll, counts, R, Rt = _estep_python(o)
    p, d = pmm_ref.RefProblem.from_oracle(o)
    cn, cc, cR, cRt = p.estep(d, 0.2, 0.1)
    assert math.isclose(cn, nll, rel_tol=1e-12)
    assert np.allclose(cc, counts, rtol=1e-9, atol=1e-10)
    assert math.isclose(cRt, Rt, rel_tol=1e-10)
    # the M-step from the same sums
    d1, nu1, phi1 = p.mstep(cc, cR, cRt, d, 0.2, 0.1)
    o.mstep()
    nodes = [v for r in o.trees for v in postorder(r)]
    want = np.asarray([v.edge_length for v in nodes])
    assert np.allclose(d1, want, rtol=1e-9)
    assert math.isclose(nu1, o.nu, rel_tol=1e-9) and math.isclose(phi1, o.phi, rel_tol=1e-9)


def test_c_em_iterations_match_python_oracle():
    T, msa, Q = load_test4()
    o = OracleEM([T], msa, Q, nu=0.3, phi=0.2)
    o.az_partition()
    p, d = pmm_ref.RefProblem.from_oracle(o)
    nll_c, it, dc, nuc, phic = p.em(d, 0.3, 0.2, fixed_iters=3)
    nll_py, it_py, _ = o.em_optimization(fixed_iters=3)
    assert it == 4 and it_py == 4
    assert math.isclose(nll_c, nll_py, rel_tol=1e-11)
    nodes = [v for r in o.trees for v in postorder(r)]
    assert np.allclose(dc, [v.edge_length for v in nodes], rtol=1e-8)
    assert math.isclose(nuc, o.nu, rel_tol=1e-8) and math.isclose(phic, o.phi, rel_tol=1e-8, abs_tol=1e-14)


def test_c_em_reaches_reference_kat_test47():
    em = KATS["em_kat"]
    T, msa, Q = load_test4()
    o = OracleEM([T], msa, Q, 0, 0)
    rng = random.Random(em["random_seed"])
    seed = int(rng.random() * 10000)
    x0 = o.ini_all(random.Random(seed))
    o.x2params(x0)
    p, d = pmm_ref.RefProblem.from_oracle(o)
    nll, it, dc, nu, phi = p.em(d, o.nu, o.phi)
    assert abs(nll - em["nllh"]) / em["nllh"] < em["nllh_rel_tol"]
    assert round(abs(phi - em["phi"]), em["param_places"]) == 0
    assert round(abs(nu - em["nu"]), em["param_places"]) == 0


def test_c_oracle_is_deterministic():
    T, msa, Q = load_example1()
    o = OracleEM([T], msa, Q, nu=0.2, phi=0.1)
    p, d = pmm_ref.RefProblem.from_oracle(o)
    n1, c1, _, _ = p.estep(d, 0.2, 0.1)
    assert pmm_ref.threads() >= 1
    n2, c2, _, _ = p.estep(d, 0.2, 0.1)
    assert n1 == n2 and np.array_equal(c1, c2)
