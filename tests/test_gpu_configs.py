"""The benchmark configurations at their full sizes, through the C ABI.

* 256 x 256 x 100 frames, ARPS on, patch 4, trajectory 15, PGURE lambda (the
  headline workload): the whole sequence denoised once, then the first,
  middle and last windows checked against the CPU FP64 oracle at the GPU's
  lambda (trajectories and noise estimates bit-exact, frames within 1e-5),
  the middle window's risk and lambda search too.
* A draw (seed 3) from the batched uneven workload: 512 x 512, 128 x 128 and
  256 x 256 sequences with patches 4, 6, 8 and trajectories 7, 15, 31,
  64 lambda candidates each, through pgsvt_denoise_batch: every sampled
  window's lambda is the least-risk candidate and its frame the oracle's at
  that lambda.
"""

import sys

import numpy as np
import pytest

import oracle_lib as O
import pgsvt_lib as G
from pguresvt import batch as B
from pguresvt import learned as Lr
from pguresvt import phantom

pytestmark = pytest.mark.gpu

FRAME_TOL = 1e-5
RISK_TOL = 1e-4


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def window_of(t, nf, nw):
    fw = nw // 2
    if t < fw:
        return 0, t
    if t >= nf - fw:
        return nf - nw, t - (nf - nw)
    return t - fw, fw


@pytest.fixture(scope="module")
def lapack():
    try:
        O.set_svd(1)  # the oracle's SVDs through LAPACK dgesdd: faster at these sizes
    except RuntimeError:
        pass
    yield
    O.set_svd(0)


def test_headline_sequence_256x256x100(gpu, lapack):
    _, X = phantom.noisy_sequence(256, 100, blobs=8, drift=2, alpha=1.0, mu=0.0, sigma=8.0, seed=2)
    kw = dict(motion_filter=-1, random_seed=2)
    log("denoise 256x256x100 on the GPU")
    Yg, eg = G.denoise(X, **kw)
    assert np.isfinite(Yg).all()
    for t in (0, 50, 99):
        s0, ref = window_of(t, 100, 15)
        log("  window t=%d" % t)
        assert np.array_equal(G.window_trajectories(X, t, **kw), O.window_trajectories(X, t, **kw))
        u = X[:, :, s0:s0 + 15].astype(np.float64)
        mx = u.max()
        assert tuple(eg[t, 1:]) == O.noise_estimate(u / mx, 4)
        o = O.window_reconstruct(X, t, eg[t, 0], **kw)
        e = G.rel_l2(Yg[:, :, t], o[:, :, ref] * mx)
        log("    frame rel L2 at the GPU lambda %.3g" % e)
        assert e < FRAME_TOL
    t = 50
    lams = [eg[t, 0] / 2, eg[t, 0], 2 * eg[t, 0]]
    rg = G.window_risk(X, t, eg[t, 1:], lams, **kw)
    ro = O.window_risk(X, t, eg[t, 1:], lams, **kw)
    log("  risk rel", np.abs(rg - ro) / np.abs(ro))
    assert np.all(np.abs(rg - ro) <= RISK_TOL * np.abs(ro))
    _, eo = O.denoise_range(X, t, t + 1, **kw)
    log("  lambda GPU %.6g oracle %.6g" % (eg[t, 0], eo[t, 0]))
    assert abs(eg[t, 0] / eo[t, 0] - 1) < 1e-2


# jobs 0, 2, 7 and 17 of the seed-3 draw of pgure-svt_amd/batch_bench.py
BATCH_JOBS = [
    dict(n=512, frames=30, patch=4, trajectory=7, seed=1000),
    dict(n=128, frames=30, patch=4, trajectory=15, seed=1002),
    dict(n=128, frames=200, patch=8, trajectory=31, seed=1007),
    dict(n=512, frames=30, patch=6, trajectory=15, seed=1017),
]


def _candidates(X, t, n, params, nw):
    """The grid of window t: log-spaced over [0.01, 10] times 1 / median^2
    of the sampled patches' singular values (exponential weighting)."""
    F, _ = Lr.features(X, t, t + 1, **params)
    s = np.exp(F[0, :256 * nw].astype(np.float64))
    m = float(np.median(s.astype(np.float32)))
    return (1.0 / (m * m)) * 10.0 ** (-2.0 + 3.0 * np.arange(n) / (n - 1))


def test_batched_uneven_draw(gpu, lapack):
    items = []
    for j in BATCH_JOBS:
        _, X = phantom.noisy_sequence(j["n"], j["frames"], blobs=8, drift=2, alpha=1.0, mu=0.0, sigma=8.0,
                                      seed=j["seed"])
        items.append((X, dict(Lr.TRAIN_PARAMS, patch_size=j["patch"], trajectory_length=j["trajectory"],
                              random_seed=3)))
    log("batch of %d sequences, 64 candidates" % len(items))
    outs = B.denoise_batch(items, n_lambda=64, concurrency=2)
    for (X, prm), (Y, est), j in zip(items, outs, BATCH_JOBS):
        nw = Lr.window_length(j["patch"], j["trajectory"])
        nf = X.shape[2]
        t = nf // 2
        s0, ref = window_of(t, nf, nw)
        log("  job %s: window t=%d lambda %.5g" % (j, t, est[t, 0]))
        assert np.isfinite(Y).all()
        lams = _candidates(X, t, 64, prm, nw)
        k = int(np.argmin(np.abs(lams / est[t, 0] - 1)))
        assert abs(lams[k] / est[t, 0] - 1) < 1e-5, "the chosen lambda is a grid candidate"
        r = G.window_risk(X, t, est[t, 1:], lams, **prm)
        assert r[k] <= r.min() + 1e-6 * abs(r.min()) + 1e-12, (k, int(np.argmin(r)))
        o = O.window_reconstruct(X, t, float(est[t, 0]), **prm)
        mx = X[:, :, s0:s0 + nw].astype(np.float64).max()
        e = G.rel_l2(Y[:, :, t], o[:, :, ref] * mx)
        log("    frame rel L2 at the chosen lambda %.3g" % e)
        assert e < FRAME_TOL
