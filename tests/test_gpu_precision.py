"""The FP64 GPU path (pgsvt_set_precision(64)) against the CPU FP64 oracle,
and the FP32 path against it within FP32 bounds."""

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import pgsvt_lib as G
from pguresvt import _pguresvt as P
from pguresvt import phantom

pytestmark = pytest.mark.gpu


def set_precision(bits):
    L = P.load()
    L.pgsvt_set_precision.argtypes = [C.c_int]
    return L.pgsvt_set_precision(bits)


@pytest.fixture
def fp64():
    assert set_precision(64) == 0
    yield
    assert set_precision(32) == 0


@pytest.fixture(scope="module")
def seq():
    _, noisy = phantom.noisy_sequence(64, 16, blobs=6, drift=1, alpha=2.0, mu=5.0, sigma=10.0, seed=4)
    return noisy


def test_set_precision_rejects_other_widths(gpu):
    assert set_precision(16) == -1
    assert set_precision(32) == 0


@pytest.mark.parametrize("bs,step", [(4, 1), (8, 2)])
def test_fp64_singular_values_match_lapack(gpu, fp64, seq, bs, step):
    kw = dict(patch_size=bs, patch_overlap=step)
    s = G.window_singular_values(seq, 8, **kw)
    O.set_svd(1)
    try:
        ref = O.window_singular_values(seq, 8, **kw)
    finally:
        O.set_svd(0)
    assert s.shape == ref.shape
    err = (np.abs(s - ref) / ref[:, :1]).max()
    assert err < 1e-12, err


@pytest.mark.parametrize("bs,step", [(4, 1), (8, 2)])
def test_fp64_reconstruction_and_risk_match_oracle(gpu, fp64, seq, bs, step):
    kw = dict(patch_size=bs, patch_overlap=step)
    lam = 0.3
    v = G.window_reconstruct(seq, 8, lam, **kw)
    ref = O.window_reconstruct(seq, 8, lam, **kw)
    assert G.rel_l2(v, ref) < 1e-12
    noise = [0.02, 0.05, 0.03]
    r = G.window_risk(seq, 8, noise, [0.1, 0.3, 1.0], **kw)
    rr = O.window_risk(seq, 8, noise, [0.1, 0.3, 1.0], **kw)
    assert np.max(np.abs(r - rr) / np.abs(rr)) < 1e-10


def test_fp32_is_within_its_bounds_of_fp64(gpu, seq):
    kw = dict(patch_size=8, patch_overlap=2)
    s32 = G.window_singular_values(seq, 8, **kw)
    v32 = G.window_reconstruct(seq, 8, 0.3, **kw)
    assert set_precision(64) == 0
    try:
        s64 = G.window_singular_values(seq, 8, **kw)
        v64 = G.window_reconstruct(seq, 8, 0.3, **kw)
    finally:
        set_precision(32)
    assert (np.abs(s32 - s64) / s64[:, :1]).max() < 1e-5
    assert G.rel_l2(v32, v64) < 1e-5
