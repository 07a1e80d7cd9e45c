"""The noise estimator's quadtree leaf statistics on the GPU (the learned
path's) against the host estimator's (the reference path's, bit-exact
against the oracle): the same tree and leaves, robust variances bit for bit
(per-element Laplacians in the host's order, exact order statistics), robust
means equal to the rounding of their IRLS sums (workgroup order on the GPU,
even/odd accumulators on the host)."""

import numpy as np
import pytest

import pgsvt_lib as G
from pguresvt import phantom

pytestmark = pytest.mark.gpu


def _frames():
    out = []
    for n, seed in ((64, 1), (128, 2), (256, 3)):
        _, X = phantom.noisy_sequence(n, 2, blobs=6, drift=1, alpha=1.5, sigma=6.0, seed=seed)
        f = X[:, :, 1].astype(np.float64)
        out.append(("phantom %d" % n, f / f.max()))
    rng = np.random.default_rng(7)
    q = np.zeros((128, 128))
    q[:64, 64:] = 0.5
    q[64:, :64] = 0.25
    q[64:, 64:] = 0.75
    q += rng.normal(0, 0.01, q.shape) * (1 + 3 * (q > 0.3))
    out.append(("quadrants 128", q))
    out.append(("noise 100 (never split)", rng.normal(0.5, 0.05, (100, 100))))
    out.append(("flat 32", np.full((32, 32), 0.25) + rng.normal(0, 1e-3, (32, 32))))
    return out


@pytest.mark.parametrize("name,frame", _frames(), ids=lambda v: v if isinstance(v, str) else "")
def test_gpu_leaves_match_host(gpu, name, frame):
    mh, vh = G.frame_leaves(frame, False)
    mg, vg = G.frame_leaves(frame, True)
    assert len(mg) == len(mh) and len(mh) > 0, (name, len(mg), len(mh))
    assert np.array_equal(vg, vh), (name, np.abs(vg - vh).max())
    d = np.abs(mg - mh)
    # the IRLS stops when its step falls below 1e-6: sums that round
    # differently can end it one step apart
    assert np.all(d <= 1e-6 * np.maximum(1.0, np.abs(mh))), (name, d.max())
    assert np.median(d / np.maximum(np.abs(mh), 1e-300)) <= 1e-12, name


def test_gpu_leaves_reject_small_frames(gpu):
    with pytest.raises(RuntimeError, match="16 x 16"):
        G.frame_leaves(np.ones((8, 8)), True)
