"""Window ranges and the learned-lambda path through the C-ABI."""

import numpy as np
import pytest

from pguresvt import learned as Lr
from pguresvt import phantom

import pgsvt_lib as G

pytestmark = pytest.mark.gpu

PARAMS = dict(Lr.TRAIN_PARAMS)


@pytest.fixture(scope="module")
def seq():
    _, noisy = phantom.noisy_sequence(64, 20, blobs=6, drift=1, alpha=1.0, sigma=6.0, seed=31)
    return noisy


@pytest.fixture
def no_probe():
    Lr.set_fallback(Lr.FALLBACK_DELTA, 0.0)
    yield
    Lr.set_fallback()


def _random_model(n_in, y_mean, seed=0):
    rng = np.random.default_rng(seed)
    return Lr.LambdaModel(Lr.init_params(n_in, 32, 1) * 0.05, rng.standard_normal(n_in), np.full(n_in, 3.0),
                          y_mean, 0.3, hidden=32)


def test_range_equals_full_call(gpu, seq):
    # windows are independent: any sub-range gives the full call's frames
    Yf, ef = Lr.denoise_range(seq, 0, 20, **PARAMS)
    for tb, te in ((0, 3), (5, 12), (17, 20)):
        Y, e = Lr.denoise_range(seq, tb, te, **PARAMS)
        assert np.array_equal(Y, Yf[:, :, tb:te])
        assert np.array_equal(e, ef[tb:te])


def test_range_rejects_bad_bounds(gpu, seq):
    with pytest.raises(RuntimeError, match="t_begin < t_end"):
        Lr.denoise_range(seq, 5, 5, **PARAMS)


def test_features_are_sampled_log_singular_values(gpu, seq):
    F, est = Lr.features(seq, 9, 10, **PARAMS)
    assert F.shape == (1, Lr.n_features(15))
    S = G.window_singular_values(seq, 9, **PARAMS)  # (patches, 15), descending
    logS = np.log(np.maximum(S, 1e-6))
    rows = F[0, :256 * 15].reshape(256, 15).astype(np.float64)
    idx = []
    for r in rows:  # the patch whose log spectrum the row is (to float32 log rounding)
        d = np.abs(logS - r).max(1)
        i = int(np.argmin(d))
        assert d[i] < 1e-5, "every feature row is the log spectrum of some patch"
        idx.append(i)
    assert idx == sorted(idx) and len(set(idx)) == 256, "256 distinct patches in index order"


def test_feature_noise_equals_driver_noise_when_window_is_sequence(gpu):
    # with as many frames as the window, the sequence maximum is the window
    # maximum: the learned path's noise features equal the driver's estimates
    _, X = phantom.noisy_sequence(64, 15, blobs=5, drift=1, alpha=2.0, sigma=4.0, seed=5)
    F, _ = Lr.features(X, 7, 8, **PARAMS)
    _, est = Lr.denoise_range(X, 7, 8, **PARAMS)
    assert np.array_equal(F[0, -3:], est[0, 1:].astype(np.float32))


def test_learned_path_applies_the_proposal(gpu, seq, no_probe):
    # a random model, no probe: each output frame equals the fixed-lambda
    # driver run at the proposed lambda of its window
    n_in = Lr.n_features(15)
    model = _random_model(n_in, np.log(0.5))
    Y, est = Lr.denoise_learned(seq, model, **PARAMS)
    assert Lr.last_fallbacks() == 0
    F, _ = Lr.features(seq, 0, 20, **PARAMS)
    lam = np.exp(model.predict_log_lambda(F))
    assert np.allclose(est[:, 0], lam, rtol=1e-5)
    for t in (0, 9, 19):
        kw = dict(PARAMS, optimize_pgure=0, lambda1=float(est[t, 0]))
        Yt, _ = Lr.denoise_range(seq, t, t + 1, **kw)
        assert np.array_equal(Y[:, :, t], Yt[:, :, 0])


def test_learned_rejects_mismatched_model(gpu, seq):
    model = Lr.LambdaModel(Lr.init_params(10, 16, 0), np.zeros(10), np.ones(10), 0.0, 1.0, hidden=16)
    with pytest.raises(RuntimeError, match="expects 10 features"):
        Lr.denoise_learned(seq, model, **PARAMS)


def _model_or_skip():
    import os

    if not os.path.exists(Lr.default_model_path()):
        pytest.skip("no trained lambda model in the package")
    return Lr.LambdaModel.load(Lr.default_model_path())


def test_trained_model_gains_over_no_network(gpu):
    # Held-out sequences (seeds outside the training range): the PGURE risk
    # of the network's lambda is closer to the searched lambda's than that of
    # the two proposals available without the network -- the search's own
    # start point (mean of the normalised window) and the training-set
    # median lambda.
    model = _model_or_skip()
    gap = {"net": [], "start": [], "median": []}
    for seed in (900001, 900002, 900003):
        X = Lr.training_sequence(seed)
        for t in (6, 15, 24):
            _, est = Lr.denoise_range(X, t, t + 1, **PARAMS)
            F, _ = Lr.features(X, t, t + 1, **PARAMS)
            lam_star, noise = est[0, 0], est[0, 1:]
            s0 = min(max(t - 7, 0), X.shape[2] - 15)  # window start (src/pguresvt.hpp:96-114)
            w = X[:, :, s0:s0 + 15].astype(np.float64)
            lam_start = (w / w.max()).mean()
            lams = [lam_star, float(np.exp(model.predict_log_lambda(F)[0])), lam_start, float(np.exp(model.y_mean))]
            r = Lr.window_risk(X, t, noise, lams, **PARAMS)
            for k, v in zip(("net", "start", "median"), r[1:]):
                gap[k].append((v - r[0]) / abs(r[0]))
    g = {k: float(np.mean(v)) for k, v in gap.items()}
    assert g["net"] < 0.5 * g["median"] and g["net"] < 0.2 * g["start"], g


def test_probe_sends_poor_proposals_to_the_search(gpu, seq, no_probe):
    # A model that proposes lambdas 30x too large: with the probe on (the
    # library's default), the windows it sends to the search come back with
    # the search's lambda, whose PGURE risk (same noise estimates) is below
    # the proposal's, and their frames are the fixed-lambda driver's at that
    # lambda; the windows it keeps carry the proposal unchanged.
    n_in = Lr.n_features(15)
    _, est0 = Lr.denoise_range(seq, 0, 20, **PARAMS)
    model = _random_model(n_in, float(np.log(30 * np.median(est0[:, 0]))))
    _, prop = Lr.denoise_learned(seq, model, **PARAMS)  # probe off: the proposals
    Lr.set_fallback(refine=0)  # no refinement: every failing window is searched
    Y, est = Lr.denoise_learned(seq, model, **PARAMS)
    n_fb = Lr.last_fallbacks()
    moved = np.flatnonzero(est[:, 0] != prop[:, 0])
    assert 0 < n_fb and len(moved) <= n_fb, (n_fb, moved)
    assert np.array_equal(est[:, 1:], prop[:, 1:])
    for t in moved[:4]:
        r = Lr.window_risk(seq, int(t), est[t, 1:], [est[t, 0], prop[t, 0]], **PARAMS)
        assert r[0] < r[1], (t, r)
        Yt, _ = Lr.denoise_range(seq, int(t), int(t) + 1, **dict(PARAMS, optimize_pgure=0, lambda1=float(est[t, 0])))
        assert np.array_equal(Y[:, :, t], Yt[:, :, 0])


def test_probe_keeps_near_optimal_proposals(gpu, seq, no_probe):
    # The searched lambdas themselves as proposals (a model that returns
    # each window's search result is emulated by a fixed lambda equal to the
    # median of the searched ones on a stationary sequence): the probe keeps
    # most windows, and with the limit raised to 1e9 it keeps every one.
    n_in = Lr.n_features(15)
    _, est0 = Lr.denoise_range(seq, 0, 20, **PARAMS)
    model = _random_model(n_in, float(np.log(np.median(est0[:, 0]))))
    model.params[:] = 0.0  # the network outputs 0: lambda = exp(y_mean) everywhere
    Lr.set_fallback(0.1, 1e9)
    Lr.denoise_learned(seq, model, **PARAMS)
    assert Lr.last_fallbacks() == 0
    Lr.set_fallback()
    _, est = Lr.denoise_learned(seq, model, **PARAMS)
    assert Lr.last_fallbacks() <= 10, Lr.last_fallbacks()


def test_probe_refines_poor_proposals(gpu, seq, no_probe):
    # With refinement (the default), a window the probe fails moves to the
    # probe's estimate of the minimum: every window whose lambda changed ends
    # with a lower PGURE risk than its proposal's, and its frame is the
    # fixed-lambda driver's at its final lambda.
    n_in = Lr.n_features(15)
    _, est0 = Lr.denoise_range(seq, 0, 20, **PARAMS)
    model = _random_model(n_in, float(np.log(30 * np.median(est0[:, 0]))))
    _, prop = Lr.denoise_learned(seq, model, **PARAMS)
    Lr.set_fallback()
    Y, est = Lr.denoise_learned(seq, model, **PARAMS)
    moved = np.flatnonzero(est[:, 0] != prop[:, 0])
    assert len(moved) > 0
    for t in moved[:6]:
        r = Lr.window_risk(seq, int(t), est[t, 1:], [est[t, 0], prop[t, 0]], **PARAMS)
        assert r[0] < r[1], (t, r)
    t = int(moved[0])
    Yt, _ = Lr.denoise_range(seq, t, t + 1, **dict(PARAMS, optimize_pgure=0, lambda1=float(est[t, 0])))
    assert np.array_equal(Y[:, :, t], Yt[:, :, 0])


def test_fallback_setter_rejects_bad_values(gpu):
    with pytest.raises(RuntimeError, match="delta"):
        Lr.set_fallback(0.0, 0.01)
    with pytest.raises(RuntimeError, match="finite"):
        Lr.set_fallback(0.1, float("nan"))
    with pytest.raises(RuntimeError, match="refinements"):
        Lr.set_fallback(0.1, 0.01, 17)
    Lr.set_fallback()
