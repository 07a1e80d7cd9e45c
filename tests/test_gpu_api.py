"""The reference test suite's SVT scenarios (pguresvt/tests/test_svt.py:61-106)
run through the Python API on the GPU."""

import os

import numpy as np
import pytest

from pguresvt import SVT

pytestmark = pytest.mark.gpu


def _nsed(A, B):
    A_m = A - A.mean()
    B_m = B - B.mean()
    return 0.5 * np.linalg.norm(A_m - B_m) ** 2 / (np.linalg.norm(A_m) ** 2 + np.linalg.norm(B_m) ** 2)


@pytest.fixture(scope="module")
def data():
    X = np.load(os.path.join(os.path.dirname(__file__), "golden", "data.npz"))["a"]
    rng = np.random.RandomState(101)
    Y = X + 100.0 + 100.0 * rng.randn(*X.shape)
    Y[Y < 0.0] = 0.0
    return X, Y.astype(np.uint16)


def test_single_threaded(gpu, data):
    X, Y = data
    s = SVT(n_jobs=1, random_seed=101)
    s.denoise(Y)
    for a in ("Y_", "lambda1s_", "noise_alphas_", "noise_mus_", "noise_sigmas_"):
        assert hasattr(s, a)
    assert s.Y_.shape == X.shape
    assert _nsed(X, s.Y_) < 0.025


def test_multi_threaded(gpu, data):
    X, Y = data
    s = SVT(random_seed=101)
    s.denoise(Y)
    assert _nsed(X, s.Y_) < 0.025


@pytest.mark.parametrize("opt", [True, False])
def test_opt_lambda1(gpu, data, opt):
    X, Y = data
    s = SVT(lambda1=5.0, optimize_pgure=opt, random_seed=101)
    s.denoise(np.asfortranarray(Y))
    assert _nsed(X, s.Y_) < 0.025


def test_fortran_array(gpu, data):
    X, Y = data
    s = SVT(random_seed=101)
    s.denoise(np.asfortranarray(Y))
    assert _nsed(X, s.Y_) < 0.025


def test_known_noise(gpu, data):
    X, Y = data
    s = SVT(noise_mu=100.0, noise_sigma=100.0, random_seed=101)
    s.denoise(Y)
    assert _nsed(X, s.Y_) < 0.3


@pytest.mark.parametrize("dtype", [np.uint8, np.float32, np.float64])
def test_dtypes(gpu, data, dtype):
    X, Y = data
    Z = (Y // 256).astype(dtype) if dtype == np.uint8 else Y.astype(dtype)
    s = SVT(random_seed=101, motion_filter=None)
    s.denoise(Z)
    assert s.Y_.shape == X.shape and np.all(np.isfinite(s.Y_))
