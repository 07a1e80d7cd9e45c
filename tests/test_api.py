"""CPU tests of the Python API, the C ABI library and the host noise estimator."""

import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
import pgsvt_lib as G
from pguresvt import SVT, mixed_noise_model, phantom
from pguresvt import _pguresvt as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _nsed(A, B):
    A_m = A - A.mean()
    B_m = B - B.mean()
    return 0.5 * np.linalg.norm(A_m - B_m) ** 2 / (np.linalg.norm(A_m) ** 2 + np.linalg.norm(B_m) ** 2)


class TestMixedNoiseModel:
    """Golden values of pguresvt/tests/test_svt.py:23-58."""

    def setup_method(self, method):
        self.rng = np.random.RandomState(101)
        self.X = self.rng.uniform(low=0, high=255, size=(64, 64, 32))

    def test_default(self):
        Y = mixed_noise_model(self.X, random_state=self.rng)
        np.testing.assert_allclose(_nsed(self.X, Y), 0.3753506, rtol=1e-6)

    def test_alpha(self):
        Y = mixed_noise_model(self.X, alpha=1e-5, random_state=self.rng)
        np.testing.assert_allclose(_nsed(self.X, Y), 1.4986839425e-05, rtol=1e-6)

    def test_mu(self):
        Y = mixed_noise_model(self.X, alpha=1e-5, mu=0.1, random_state=self.rng)
        np.testing.assert_allclose(_nsed(self.X, Y), 1.4986839425e-05, rtol=1e-6)
        np.testing.assert_allclose(self.X.mean(), Y.mean() - 0.1 * self.X.max(), rtol=5e-5)

    def test_sigma(self):
        Y = mixed_noise_model(self.X, alpha=1e-5, sigma=0.1, random_state=self.rng)
        np.testing.assert_allclose(_nsed(self.X, Y), 0.02836968, rtol=1e-6)

    def test_error_alpha(self):
        with pytest.raises(ValueError, match="alpha should be in range"):
            mixed_noise_model(self.X, alpha=-1.0)

    def test_error_sigma(self):
        with pytest.raises(ValueError, match="sigma should be"):
            mixed_noise_model(self.X, sigma=-1.0)

    @pytest.mark.parametrize("rng", [None, 101, np.random.RandomState(101)])
    def test_random_state(self, rng):
        mixed_noise_model(self.X, random_state=rng)


class TestErrors:
    """Validation messages of pguresvt/tests/test_svt.py:109-164."""

    def setup_method(self, method):
        self.X = np.ones((32, 32, 16))

    def test_negative_data(self):
        with pytest.raises(ValueError, match="Negative values found in data"):
            self.X[0, 0, 0] = -1
            SVT().denoise(self.X)

    def test_error_patch_overlap(self):
        with pytest.raises(ValueError, match="Invalid patch_overlap parameter"):
            SVT(patch_size=10, patch_overlap=11).denoise(self.X)

    def test_error_trajectory_length(self):
        with pytest.raises(ValueError, match="Invalid trajectory_length parameter"):
            SVT(trajectory_length=14).denoise(self.X)

    def test_error_motion_window(self):
        with pytest.raises(ValueError, match="Invalid motion_window parameter"):
            SVT(motion_estimation=True, motion_window=1).denoise(self.X)
        with pytest.raises(ValueError, match="Invalid motion_window parameter"):
            SVT(motion_estimation=True, motion_window=8).denoise(self.X)

    def test_error_motion_filter(self):
        with pytest.raises(ValueError, match="Invalid motion_filter parameter"):
            SVT(motion_estimation=True, motion_filter=0.5).denoise(self.X)

    def test_error_lambda1(self):
        with pytest.raises(ValueError, match="Invalid lambda1 parameter"):
            SVT(optimize_pgure=False, lambda1=-1.0).denoise(self.X)

    def test_error_non_square(self):
        m, _, _ = self.X.shape
        with pytest.raises(ValueError, match="requires square images"):
            SVT().denoise(self.X[: m // 2, :, :])

    def test_error_non_power_of_two(self):
        m, n, _ = self.X.shape
        with pytest.raises(ValueError, match="requires image dimensions 2\\^N"):
            SVT().denoise(self.X[: m - 1, : n - 1, :])

    def test_error_dtype(self):
        with pytest.raises(TypeError, match="Invalid dtype"):
            SVT().denoise(self.X == 0)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "pgsvt.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(pgsvt_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = P.load()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n


def test_library_reports_version_and_defaults():
    L = P.load()
    assert L.pgsvt_version().startswith(b"pgsvt")
    p = P.pgsvt_params()
    L.pgsvt_default_params(C.byref(p))
    assert (p.trajectory_length, p.patch_size, p.patch_overlap, p.motion_window, p.motion_filter) == (15, 4, 1, 7, 5)
    assert (p.noise_method, p.max_iter, p.n_jobs, p.random_seed) == (4, 500, -1, -1)
    assert p.optimize_pgure == 1 and p.exponential_weighting == 1 and p.motion_estimation == 1
    assert p.tol == 1e-7


def test_library_rejects_bad_arguments_without_gpu():
    X = np.ones((32, 16, 16), np.float32, order="F")  # non-square
    p = G.params()
    Y = np.zeros(X.shape, order="F")
    E = np.zeros((16, 4), order="F")
    rc = P.load().pgsvt_denoise(X.ctypes.data, 2, 32, 16, 16, C.byref(p), Y.ctypes.data, E.ctypes.data)
    assert rc == -1 and b"square" in P.load().pgsvt_last_error()
    Xs = np.ones((32, 32, 10), np.float32, order="F")  # fewer frames than the window
    rc = P.load().pgsvt_denoise(Xs.ctypes.data, 2, 32, 32, 10, C.byref(p), Y.ctypes.data, E.ctypes.data)
    assert rc == -1


@pytest.mark.parametrize("seed,n,t0", [(2, 64, 5), (11, 32, 0), (5, 128, 3)])
def test_noise_estimator_matches_oracle(seed, n, t0):
    # the product's optimised host estimator is the same arithmetic as the
    # oracle's straight restatement: results are bit-identical
    clean, noisy = phantom.noisy_sequence(n, 20, drift=2, alpha=1.0, sigma=8.0, seed=seed)
    u = noisy[:, :, t0:t0 + 15].astype(np.float64)
    u /= u.max()
    for method in (4, 3, 2, 1):
        g = G.noise_estimate(u, method)
        o = O.noise_estimate(u, method)
        assert g == o, (method, g, o)
    g = G.noise_estimate(u, 4, alpha=0.5, mu=-1.0, sigma=0.1)
    o = O.noise_estimate(u, 4, alpha=0.5, mu=-1.0, sigma=0.1)
    assert g == o and g[0] == 0.5 and g[2] == 0.1


@pytest.mark.parametrize("levels", [7, 255])
def test_noise_estimator_matches_oracle_on_quantised_frames(levels):
    # integer-valued frames put long runs of equal values into every order
    # statistic and histogram bucket of the estimator's selections, and flat
    # clipped regions give leaves of a single value
    clean, noisy = phantom.noisy_sequence(128, 15, drift=1, alpha=1.0, sigma=8.0, seed=3)
    q = np.clip(np.round(noisy * (levels / noisy.max())), 0, levels - 2)
    u = q.astype(np.float64) / q.max()
    for method in (4, 2):
        assert G.noise_estimate(u, method) == O.noise_estimate(u, method), method


def test_noise_estimator_recovers_known_line():
    # a robust WLS line fit on exact (mean, var) pairs must recover the line:
    # pure Gaussian noise of std s on a flat frame gives alpha ~ 0 and sigma ~ s
    rng = np.random.default_rng(0)
    u = 0.3 + 0.01 * rng.standard_normal((64, 64, 5))
    a, m, s = G.noise_estimate(u, 4)
    assert abs(m - 0.3) < 2e-3
    assert abs(s - 0.01) < 2e-3


def test_noise_method1_too_few_leaves_is_numeric_error():
    # an all-zero window has one quadtree leaf per frame, fewer than noise
    # method 1 reads (round(0.05 * 4369) + 1 = 219 of 15); the reference
    # aborts in its bounds-checked copy (src/noise.hpp:115-121, 384-392), the
    # library returns PGSVT_E_NUMERIC with the counts in the message
    u = np.zeros((256, 256, 15))
    with pytest.raises(P.PGSVTError, match="error -3.*219 quadtree leaves.*15"):
        G.noise_estimate(u, 1)
    # the other methods need no such count
    for method in (2, 3, 4):
        assert G.noise_estimate(u, method) == O.noise_estimate(u, method)


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("kind", ["flat", "textured"])
def test_noise_methods_1_to_3_match_oracle(method, kind):
    # methods 1-3 (the modes of the leaf statistics, src/noise.hpp:115-147)
    # on a flat noisy frame (few, large leaves) and a textured one (many
    # small leaves): the product's estimator equals the oracle bit for bit
    rng = np.random.default_rng(7 + method)
    if kind == "flat":
        u = 0.4 + 0.02 * rng.standard_normal((64, 64, 9))
    else:
        clean, noisy = phantom.noisy_sequence(64, 9, drift=1, alpha=0.5, sigma=5.0, seed=method)
        u = noisy.astype(np.float64)
    u = u / u.max()
    if kind == "flat" and method == 1:
        # a flat frame is one leaf, 9 leaves against the 24 method 1 reads:
        # both refuse (the reference aborts)
        with pytest.raises(P.PGSVTError, match="error -3"):
            G.noise_estimate(u, method)
        with pytest.raises(RuntimeError):
            O.noise_estimate(u, method)
        return
    assert G.noise_estimate(u, method) == O.noise_estimate(u, method)
