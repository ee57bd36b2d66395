"""Pins of the FP64 restatement of the surrogate network (oracle/mlp_ref.py),
the checker of the BF16-MFMA kernels: its gradients against central finite
differences of its loss, a one-layer linear network against the closed-form
least-squares gradient, GELU's known values and its agreement with the erf
form, and bfloat16 rounding of known bit patterns."""
import math

import numpy as np
import pytest

from oracle import mlp_ref


def small_net(dims, seed):
    rng = np.random.default_rng(seed)
    ws = [rng.normal(0, 1 / math.sqrt(dims[l]), size=(dims[l + 1], dims[l]))
          for l in range(len(dims) - 1)]
    bs = [rng.normal(0, 0.1, size=dims[l + 1]) for l in range(len(dims) - 1)]
    return ws, bs


@pytest.mark.parametrize("dims", [[6, 5, 4, 3], [9, 7, 7, 7, 5]])
def test_gradients_match_finite_differences(dims):
    ws, bs = small_net(dims, 1)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(11, dims[0]))
    t = rng.normal(size=(11, dims[-1]))
    wout = rng.uniform(0.5, 2.0, size=dims[-1])
    L, gws, gbs = mlp_ref.gradients(ws, bs, x, t, wout)
    assert abs(L - mlp_ref.loss(ws, bs, x, t, wout)) <= 1e-14 * L
    h = 1e-6
    for l in range(len(ws)):
        for (a, b) in [(0, 0), (dims[l + 1] - 1, dims[l] - 1), (1, 2)]:
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[l][a, b] += h
            wm[l][a, b] -= h
            fd = (mlp_ref.loss(wp, bs, x, t, wout) - mlp_ref.loss(wm, bs, x, t, wout)) / (2 * h)
            assert abs(fd - gws[l][a, b]) <= 1e-7 * max(1.0, abs(fd)), (l, a, b)
        for a in (0, dims[l + 1] - 1):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][a] += h
            bm[l][a] -= h
            fd = (mlp_ref.loss(ws, bp, x, t, wout) - mlp_ref.loss(ws, bm, x, t, wout)) / (2 * h)
            assert abs(fd - gbs[l][a]) <= 1e-7 * max(1.0, abs(fd)), (l, a)


def test_linear_layer_closed_form():
    """One linear layer: y = W x + b, loss = sum (y - t)^2 / (R N), so
    dW = 2 (y - t)^T x / (R N) and db = 2 sum_r (y - t) / (R N)."""
    rng = np.random.default_rng(3)
    W, b = rng.normal(size=(4, 6)), rng.normal(size=4)
    x, t = rng.normal(size=(9, 6)), rng.normal(size=(9, 4))
    L, gws, gbs = mlp_ref.gradients([W], [b], x, t)
    r = x @ W.T + b - t
    assert np.allclose(L, np.sum(r * r) / 36)
    assert np.allclose(gws[0], 2 * r.T @ x / 36, rtol=1e-13, atol=0)
    assert np.allclose(gbs[0], 2 * r.sum(0) / 36, rtol=1e-13, atol=0)


def test_gelu_values():
    assert mlp_ref.gelu(0.0) == 0.0
    assert abs(mlp_ref.gelu(10.0) - 10.0) < 1e-12
    assert abs(mlp_ref.gelu(-10.0)) < 1e-12
    assert abs(mlp_ref.gelu_grad(0.0) - 0.5) < 1e-15
    z = np.linspace(-6, 6, 241)
    exact = 0.5 * z * (1 + np.array([math.erf(v / math.sqrt(2)) for v in z]))
    assert np.max(np.abs(mlp_ref.gelu(z) - exact)) < 1e-3
    h = 1e-6
    fd = (mlp_ref.gelu(z + h) - mlp_ref.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - mlp_ref.gelu_grad(z))) < 1e-8


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1e-30, 3.0e38])
    r = mlp_ref.bf16_round(v)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # tie rounds to the even mantissa
    assert r[2] == 1.0 + 2 ** -7       # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 1e-30) <= 1e-30 * 2 ** -8
    assert abs(r[5] - 3.0e38) <= 3.0e38 * 2 ** -8
