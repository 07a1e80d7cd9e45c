"""PGURE over a lambda grid in captured graphs, and batches of sequences."""

import os

import numpy as np
import pytest

from pguresvt import batch as B
from pguresvt import learned as Lr
from pguresvt import phantom

import pgsvt_lib as G

pytestmark = pytest.mark.gpu

PARAMS = dict(Lr.TRAIN_PARAMS)


@pytest.fixture(scope="module")
def seq():
    _, noisy = phantom.noisy_sequence(64, 20, blobs=6, drift=1, alpha=1.0, sigma=6.0, seed=77)
    return noisy


def _grid(X, t, n, params):
    """The candidates of window t: log-spaced over [0.01, 10] times 1 /
    median^2 of the sampled patches' singular values (exponential
    weighting), the sample being the learned-path features' patches."""
    F, _ = Lr.features(X, t, t + 1, **params)
    s = np.exp(F[0, :256 * 15].astype(np.float64))
    m = float(np.median(s.astype(np.float32)))
    return (1.0 / (m * m)) * 10.0 ** (-2.0 + 3.0 * np.arange(n) / (n - 1))


def test_grid_choice_is_the_least_risk_candidate(gpu, seq):
    n = 16
    Y, est = B.denoise_grid(seq, n, **PARAMS)
    for t in (0, 9, 19):
        lams = _grid(seq, t, n, PARAMS)
        j = int(np.argmin(np.abs(lams / est[t, 0] - 1)))
        assert abs(lams[j] / est[t, 0] - 1) < 1e-5, "the chosen lambda is a grid candidate"
        r = G.window_risk(seq, t, est[t, 1:], lams, **PARAMS)
        assert r[j] <= r.min() + 1e-6 * abs(r.min()) + 1e-12, (t, j, int(np.argmin(r)), r)
        # the frame is the fixed-lambda driver's at that lambda (up to the
        # autotuned SVD mapping's rounding)
        Yt, _ = Lr.denoise_range(seq, t, t + 1, **dict(PARAMS, optimize_pgure=0, lambda1=float(est[t, 0])))
        assert G.rel_l2(Y[:, :, t], Yt[:, :, 0]) < 1e-5


def test_graph_replays_match_one_batch(gpu, seq, monkeypatch):
    # 3 batches (7 + 7 + 6 windows): the graph is replayed, then re-captured
    # for the short batch; every window must come out as in one batch of 20
    Y1, e1 = B.denoise_grid(seq, 8, **PARAMS)
    monkeypatch.setenv("PGSVT_BATCH", "7")
    Y3, e3 = B.denoise_grid(seq, 8, **PARAMS)
    assert np.array_equal(e1, e3)
    assert np.array_equal(Y1, Y3)


def test_batch_of_uneven_sequences(gpu):
    items = []
    for k, (n, f, p, t) in enumerate([(64, 20, 4, 15), (32, 16, 4, 7), (64, 18, 8, 7), (32, 20, 3, 9)]):
        _, X = phantom.noisy_sequence(n, f, blobs=5, drift=1, alpha=1.5, sigma=5.0, seed=300 + k)
        items.append((X, dict(PARAMS, patch_size=p, trajectory_length=t)))
    outs = B.denoise_batch(items, n_lambda=8, concurrency=3)
    for (X, prm), (Y, est) in zip(items, outs):
        Ys, es = B.denoise_grid(X, 8, **prm)
        assert np.array_equal(es, est)
        assert np.array_equal(Ys, Y)


def test_batch_reports_the_failing_job(gpu):
    _, X = phantom.noisy_sequence(32, 16, blobs=4, seed=5)
    bad = (X[:, :20, :], dict(PARAMS))  # not square
    with pytest.raises(RuntimeError, match="job 1"):
        B.denoise_batch([(X, dict(PARAMS)), bad], n_lambda=4, concurrency=2)
