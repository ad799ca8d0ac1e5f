"""Pins the oracle's FP64 network (8-64-64-64-5, SiLU), the reference for the surrogate's
gradient checks: backward against central finite differences of the forward, forward against
closed forms (bias-only network; a single active path through SiLU)."""
import math

import numpy as np

N_PARAMS = 64 * 8 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 5 * 64 + 5


def test_param_count(O):
    assert O.lib().or_mlp_n_params() == N_PARAMS == 9221


def test_forward_closed_forms(O):
    P = np.zeros(N_PARAMS)
    b4 = P[-5:]
    b4[:] = [0.1, -0.2, 0.3, 0.4, -0.5]
    x = np.random.default_rng(0).normal(size=(7, 8))
    np.testing.assert_array_equal(O.mlp_forward(P, x), np.tile(b4, (7, 1)))
    # one path: x0 -> h1[0] -> h2[0] -> h3[0] -> y[2], unit weights
    P[:] = 0
    P[0] = 1.0                                  # W1[0][0]
    o2 = 64 * 8 + 64
    P[o2] = 1.0                                 # W2[0][0]
    o3 = o2 + 64 * 64 + 64
    P[o3] = 1.0                                 # W3[0][0]
    o4 = o3 + 64 * 64 + 64
    P[o4 + 2 * 64] = 1.0                        # W4[2][0]
    silu = lambda v: v / (1 + math.exp(-v))
    for x0 in (-2.0, 0.3, 1.7):
        xx = np.zeros((1, 8))
        xx[0, 0] = x0
        y = O.mlp_forward(P, xx)[0]
        assert abs(y[2] - silu(silu(silu(x0)))) < 1e-15
        assert np.all(y[[0, 1, 3, 4]] == 0)


def test_backward_finite_differences(O):
    rng = np.random.default_rng(1)
    P = rng.normal(0, 0.3, N_PARAMS)
    x = rng.normal(size=(16, 8))
    y = rng.normal(size=(16, 5))
    loss, g = O.mlp_loss_grad(P, x, y)
    pred = O.mlp_forward(P, x)
    assert abs(loss - ((pred - y) ** 2).mean()) < 1e-14 * loss
    idx = rng.choice(N_PARAMS, 120, replace=False)
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (O.mlp_loss_grad(Pp, x, y)[0] - O.mlp_loss_grad(Pm, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])
