"""The hand-traced ARPS cases and the single-patch threshold closed form
(tests/test_oracle_pins.py) through the GPU library's C ABI: the same
expected positions bit for bit, the same closed form within the frame
tolerance."""

import numpy as np
import pytest

import pgsvt_lib as G
import test_oracle_pins as T

pytestmark = pytest.mark.gpu


def _gpu_pos(src, d0, d2):
    X = np.asfortranarray(np.stack([d0, src, d2], axis=2))
    tr = G.window_trajectories(X, 1, trajectory_length=3, patch_size=T.BS, motion_filter=-1,
                               motion_estimation=True, motion_window=T.MW)
    n1 = T.N12 - T.BS + 1
    return lambda k, i, j: tuple(tr[k, i + n1 * j])


def test_arps_hand_traces_on_gpu(gpu):
    src, d0, d2 = T._base(1)
    for j in (0, 4):
        d2[7:9, j:j + 2] = src[5:7, j:j + 2]
        d2[5:7, j:j + 2] = src[5:7, j:j + 2] + 0.7
    pos = _gpu_pos(src, d0, d2)
    assert pos(2, 5, 0) == (7, 0) and pos(2, 5, 4) == (5, 4)  # step 2 in column 0 only

    src, d0, d2 = T._base(2)
    src[5:7, 6:8] = [[1.0, 1.5], [2.0, 2.5]]
    d2[5:7, 8:10] = src[5:7, 6:8]
    d2[5:7, 7] = src[5:7, 6]
    d0[5:7, 4:6] = src[5:7, 6:8]
    d0[5:7, 6:8] = src[5:7, 6:8] + 0.3
    pos = _gpu_pos(src, d0, d2)
    assert pos(2, 5, 6) == (5, 8) and pos(0, 5, 6) == (5, 4)  # first backward pass predictor

    src, d0, d2 = T._base(3)
    d0[3:5, 0:2] = src[5:7, 0:2]
    d0[7:9, 0:2] = src[5:7, 0:2]
    assert _gpu_pos(src, d0, d2)(0, 5, 0) == (3, 0)  # large-diamond tie: first point

    src, d0, d2 = T._base(4)
    d2[7:11, 6:8] = 5.0
    assert _gpu_pos(src, d0, d2)(2, 8, 6) == (7, 6)  # small-diamond tie: earlier point


@pytest.mark.parametrize("exp_w", [1, 0])
def test_threshold_closed_form_on_gpu(gpu, exp_w):
    rng = np.random.default_rng(11)
    bs, nw = 4, 15
    Q1, _ = np.linalg.qr(np.column_stack([np.ones(16), rng.standard_normal((16, 14))]))
    Q2, _ = np.linalg.qr(np.column_stack([np.ones(15), rng.standard_normal((15, 14))]))
    s = np.array([400.0, 6.5, 4.0, 3.0, 2.5, 2.0, 1.6, 1.3, 1.0, 0.8, 0.6, 0.4, 0.3, 0.2, 0.1])
    A = Q1 @ np.diag(s) @ Q2.T
    X = np.asfortranarray(np.stack([A[:, k].reshape(bs, bs, order="F") for k in range(nw)], axis=2))
    su = s / X.max()
    for lam in ([50.0, 300.0, 1500.0, 8000.0] if exp_w else [0.005, 0.02, 0.1, 0.2]):
        vol = G.window_reconstruct(X, 7, lam, trajectory_length=15, patch_size=bs, exponential_weighting=bool(exp_w))
        if exp_w:
            sp = np.maximum(su - su.max() * np.exp(-0.5 * lam * su * su), 0.0)
        else:
            sp = np.maximum(su - lam, 0.0)
        B = Q1 @ np.diag(sp) @ Q2.T
        want = np.stack([B[:, k].reshape(bs, bs, order="F") for k in range(nw)], axis=2)
        assert G.rel_l2(vol, want) < 1e-5, lam
