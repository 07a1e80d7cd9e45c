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


class _Signal2D:
    """The few HyperSpy signal methods the adapter uses, for an image stack
    (navigation axis: time; data shape (t, y, x))."""

    def __init__(self, data, title):
        from types import SimpleNamespace

        self.data = data
        self.axes_manager = SimpleNamespace(signal_dimension=2)
        self.metadata = SimpleNamespace(General=SimpleNamespace(title=title))
        self.folded = True

    def unfold_navigation_space(self):
        self.folded = False

    def fold(self):
        self.folded = True

    def as_signal1D(self, spectral_axis):
        from types import SimpleNamespace

        assert spectral_axis == 0 and not self.folded
        # navigation (y, x), signal axis t last, as HyperSpy aligns it
        return SimpleNamespace(_data_aligned_with_axes=np.moveaxis(self.data, 0, -1))

    def _deepcopy_with_new_data(self, data):
        return _Signal2D(data, self.metadata.General.title)


def test_hspy_adapter(gpu, data):
    # pguresvt/hspy.py:7-80: image stacks go through SVT as (y, x, t) and come
    # back in the signal's (t, y, x) order with a "Denoised" title
    from pguresvt.hspy import HSPYSVT

    _, Y = data
    stack = np.ascontiguousarray(np.moveaxis(Y, -1, 0))  # (t, y, x)
    sig = _Signal2D(stack, "noisy stack")
    kw = dict(patch_size=4, trajectory_length=5, motion_filter=None, random_seed=7)
    out = HSPYSVT(**kw).denoise(sig)
    assert sig.folded
    assert out.metadata.General.title == "Denoised noisy stack"
    assert out.data.shape == stack.shape
    ref = SVT(**kw).denoise(np.moveaxis(stack, 0, -1)).Y_
    assert np.array_equal(out.data, np.transpose(ref, (2, 0, 1)))


def test_hspy_rejects_other_dimensions(gpu, data):
    from pguresvt.hspy import HSPYSVT

    sig = _Signal2D(np.zeros((4, 8, 8)), "x")
    sig.axes_manager.signal_dimension = 3
    with pytest.raises(NotImplementedError, match="Expected 1D or 2D signal - got dimension 3"):
        HSPYSVT().denoise(sig)
