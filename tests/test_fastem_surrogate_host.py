"""The reference's fastEM surrogate M-step (fast_em/laml.py:62-110) as the
opt-in fastEM M-step restates it, on the CPU: the loss equals a direct
evaluation of M_step_loss_fn's terms, its gradient equals central
differences, and 100 L-BFGS iterations end at a stationary point of the
free variables while the fixed ones keep their values."""
import math

import numpy as np

from laml_amd.fastem import surrogate_loss, surrogate_mstep


def _case(N=40, seed=3):
    rng = np.random.default_rng(seed)
    S = rng.uniform(0.0, 30.0, size=(N, 5))
    S[::7, 2] = 0.0
    x = np.concatenate([np.log(rng.uniform(0.01, 2.0, N)), [math.log(0.2 / 0.8), math.log(0.3 / 0.7)]])
    return S, x, 900.0, 3100.0, 420.0


def test_loss_restates_m_step_loss_fn():
    S, x, nm, nnm, rt = _case()
    N = S.shape[0]
    d = np.exp(x[:N])
    nu, phi = 0.2, 0.3
    want = 0.0
    for v in range(N):
        c1 = -S[v, 0] * d[v] * (1 + nu)
        c2 = S[v, 1] * (math.log(1 - math.exp(-d[v])) - d[v] * nu)
        c3 = S[v, 2] * math.log(1 - math.exp(-d[v] * nu))
        c4 = -S[v, 3] * d[v] * nu
        c5 = S[v, 4] * math.log(1 - math.exp(-d[v] * nu))
        want += c1 + c2 + c3 + c4 + c5
    want += nnm * math.log(1 - phi) + rt * math.log(phi)
    loss, _ = surrogate_loss(x, S, nm, nnm, rt)
    assert math.isclose(loss, -want, rel_tol=1e-12)


def test_gradient_matches_central_differences():
    S, x, nm, nnm, rt = _case()
    _, g = surrogate_loss(x, S, nm, nnm, rt)
    h = 1e-6
    for i in list(range(0, len(x) - 2, 5)) + [len(x) - 2, len(x) - 1]:
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        fd = (surrogate_loss(xp, S, nm, nnm, rt)[0] - surrogate_loss(xm, S, nm, nnm, rt)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * (1 + abs(fd)), (i, fd, g[i])


def test_mstep_reaches_stationarity_and_keeps_fixed_values():
    S, x, nm, nnm, rt = _case(N=30, seed=5)
    free = np.ones(len(x), dtype=bool)
    free[::4] = False   # some branches fixed
    free[-2] = False    # nu fixed
    xn = surrogate_mstep(x, free, S, nm, nnm, rt)
    np.testing.assert_array_equal(xn[~free], x[~free])
    loss0, _ = surrogate_loss(x, S, nm, nnm, rt)
    loss1, g = surrogate_loss(xn, S, nm, nnm, rt)
    assert loss1 < loss0
    assert np.max(np.abs(g[free])) <= 1e-6 * (1 + abs(loss1)), np.max(np.abs(g[free]))
