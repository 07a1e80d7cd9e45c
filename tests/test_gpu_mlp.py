"""The learned-lambda network on the GPU (BF16 matrix cores, FP32 sums)
against the CPU oracle oracle/mlp_ref.py, which rounds every product operand
to bfloat16 the same way and sums in FP64."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref as R  # noqa: E402

from pguresvt import learned as Lr  # noqa: E402

pytestmark = pytest.mark.gpu


def _case(n, n_in, h, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, n_in)).astype(np.float32)
    t = rng.standard_normal(n).astype(np.float32)
    p = Lr.init_params(n_in, h, seed)
    p[n_in * h:n_in * h + h] = rng.standard_normal(h) * 0.1  # non-zero biases exercise the bias paths
    return X, t, p


@pytest.mark.parametrize("n,n_in,h", [(16, 16, 16), (37, 70, 32), (5, 3843, 128), (130, 200, 128)])
def test_forward_matches_oracle(gpu, n, n_in, h):
    # a wrong operand lane layout of the 16x16x16 bf16 matrix op, a swapped
    # stride or a dropped k-tail would all show here
    X, _, p = _case(n, n_in, h, n + n_in)
    y = Lr.mlp_forward(p, X, n_in, h)
    ref, _ = R.forward(p, X.astype(np.float64), n_in, h)
    scale = np.abs(ref).max() + 1e-12
    assert np.abs(y - ref).max() / scale < 2e-3


@pytest.mark.parametrize("n,n_in,h", [(16, 16, 16), (37, 70, 32), (64, 300, 128)])
def test_gradient_matches_oracle(gpu, n, n_in, h):
    X, t, p = _case(n, n_in, h, 7 * n + h)
    g, loss = Lr.mlp_gradient(p, X, t, n_in, h)
    gr, lr = R.gradient(p, X.astype(np.float64), t.astype(np.float64), n_in, h)
    assert abs(loss - lr) <= 1e-3 * abs(lr) + 1e-6
    # per parameter block, relative to the block's largest gradient
    offs = np.cumsum([0, n_in * h, h, h * h, h, h * h, h, h, 1])
    for a, b in zip(offs[:-1], offs[1:]):
        scale = np.abs(gr[a:b]).max() + 1e-12
        assert np.abs(g[a:b] - gr[a:b]).max() / scale < 5e-3, (a, b)


def test_gradient_is_the_loss_derivative(gpu):
    # independent of the oracle: a central difference of the GPU loss along
    # the GPU gradient (bf16 rounding makes the loss piecewise smooth, so use
    # a step well above its resolution)
    X, t, p = _case(32, 48, 32, 3)
    g, _ = Lr.mlp_gradient(p, X, t, 48, 32)
    d = g / np.linalg.norm(g)
    eps = 2e-2
    _, lp = Lr.mlp_gradient(p + eps * d, X, t, 48, 32)
    _, lm = Lr.mlp_gradient(p - eps * d, X, t, 48, 32)
    slope = (lp - lm) / (2 * eps)
    assert abs(slope - np.linalg.norm(g)) < 0.1 * np.linalg.norm(g)


def test_training_fits_a_smooth_target(gpu):
    rng = np.random.default_rng(11)
    n, n_in = 2048, 24
    X = rng.standard_normal((n, n_in)).astype(np.float32)
    w = rng.standard_normal(n_in) / np.sqrt(n_in)
    t = (np.tanh(X @ w) + 0.1 * (X[:, 0] ** 2)).astype(np.float32)
    p0 = Lr.init_params(n_in, 64, 1)
    _, loss0 = Lr.mlp_gradient(p0, X, t, n_in, 64)
    p, losses = Lr.mlp_train(X, t, p0, epochs=40, batch=128, lr=2e-3, seed=2, h=64)
    _, loss1 = Lr.mlp_gradient(p, X, t, n_in, 64)
    assert losses[-1] < 0.2 * losses[0]
    assert loss1 < 0.05 * loss0
