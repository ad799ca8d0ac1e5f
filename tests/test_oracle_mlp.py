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


def bf16(a):
    """Nearest BF16 (ties to even) of the FP32 rounding of a."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_forward_bf16_roundings(O):
    """The BF16 inference restatement: on a single unit-weight path each activation is rounded to
    BF16 (through FP32, ties to even) before the next layer; with BF16-exact weights and inputs a
    generic network differs from the FP64 one only by those roundings, at most 2^-9 relative per
    activation."""
    P = np.zeros(N_PARAMS)
    P[0] = 1.0
    o2 = 64 * 8 + 64
    P[o2] = 1.0
    o3 = o2 + 64 * 64 + 64
    P[o3] = 1.0
    o4 = o3 + 64 * 64 + 64
    P[o4 + 2 * 64] = 1.0
    silu = lambda v: v / (1 + np.exp(-v))
    x = np.zeros((5, 8))
    x[:, 0] = [-2.3, 0.31, 1.7, 1.0 + 2.0 ** -8, 5.123]
    y = O.mlp_forward_bf16(P, x)
    h = bf16(x[:, 0])
    for _ in range(3):
        h = bf16(silu(h.astype(np.float32).astype(np.float64)))
    np.testing.assert_array_equal(y[:, 2], h)
    assert bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0            # a tie goes to the even neighbour
    assert bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6
    rng = np.random.default_rng(3)
    P = bf16(rng.normal(0, 0.3, N_PARAMS))
    x = bf16(rng.normal(0, 1, (200, 8)))
    y16, y64 = O.mlp_forward_bf16(P, x), O.mlp_forward(P, x)
    err = np.abs(y16 - y64).max(axis=1) / np.abs(y64).max(axis=1)
    assert 0 < err.max() < 0.05 and np.median(err) < 1e-2
