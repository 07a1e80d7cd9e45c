"""GPU path (libpgsvt.so, through the C ABI) against the CPU FP64 oracle.

Bars: ARPS trajectories and delta2 signs bit-exact; reconstructed frames
within 1e-5 relative L2; PGURE risk within 1e-4 relative; singular values
within 1e-5 of the largest one of their patch.
"""

import ctypes as C
import json
import sys

import numpy as np
import pytest

import oracle_lib as O
import pgsvt_lib as G
from pguresvt import _pguresvt as P
from pguresvt import phantom

pytestmark = pytest.mark.gpu

FRAME_TOL = 1e-5
RISK_TOL = 1e-4


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def seq64():
    clean, noisy = phantom.noisy_sequence(64, 24, drift=2, alpha=1.0, sigma=8.0, seed=2)
    return noisy


def test_delta2_signs(gpu):
    log("delta2 signs seed 101")
    n = 61440
    s = G.delta2_signs(101, n)
    d1, d2 = O.perturbations(101, n)
    assert np.array_equal(s, (d2 < 0).astype(np.uint8))


@pytest.mark.parametrize("t", [0, 3, 7, 12, 18, 21, 23])
@pytest.mark.parametrize("mf", [-1, 2, 5])  # 5: the API default median radius
def test_trajectories_bit_exact(gpu, seq64, t, mf):
    log("trajectories t=%d motion_filter=%d" % (t, mf))
    g = G.window_trajectories(seq64, t, motion_filter=mf)
    o = O.window_trajectories(seq64, t, motion_filter=mf)
    assert np.array_equal(g, o)


def test_trajectories_motion_off(gpu, seq64):
    g = G.window_trajectories(seq64, 9, motion_estimation=False)
    o = O.window_trajectories(seq64, 9, motion_estimation=0)
    assert np.array_equal(g, o)


def test_singular_values(gpu, seq64):
    log("singular values")
    g = G.window_singular_values(seq64, 10)
    o = O.window_singular_values(seq64, 10)
    assert g.shape == o.shape
    err = np.max(np.abs(g - o) / o[:, :1])
    log("  max |ds|/smax =", err)
    assert err < 1e-5


@pytest.mark.parametrize("lam,expw", [(0.3, True), (5.0, True), (60.0, True), (0.02, False), (0.2, False)])
def test_reconstruct(gpu, seq64, lam, expw):
    log("reconstruct lambda=%g exp=%s" % (lam, expw))
    g = G.window_reconstruct(seq64, 11, lam, exponential_weighting=expw)
    o = O.window_reconstruct(seq64, 11, lam, exponential_weighting=int(expw))
    e = G.rel_l2(g, o)
    log("  rel L2 =", e)
    assert e < FRAME_TOL


def test_reconstruct_motion_off(gpu, seq64):
    g = G.window_reconstruct(seq64, 2, 4.0, motion_estimation=False)
    o = O.window_reconstruct(seq64, 2, 4.0, motion_estimation=0)
    assert G.rel_l2(g, o) < FRAME_TOL


@pytest.mark.parametrize("t", [0, 12, 23])
def test_risk(gpu, seq64, t):
    log("risk t=%d" % t)
    noise = [0.012, 0.05, 0.06]
    lams = [0.5, 2.0, 10.0, 40.0]
    g = G.window_risk(seq64, t, noise, lams)
    o = O.window_risk(seq64, t, noise, lams)
    rel = np.abs(g - o) / np.abs(o)
    log("  risk rel err =", rel)
    assert np.all(rel < RISK_TOL)


def test_denoise_fixed_lambda(gpu, seq64):
    log("denoise fixed lambda")
    kw = dict(optimize_pgure=False, lambda1=3.0)
    Yg, eg = G.denoise(seq64, **kw)
    Yo, eo = O.denoise(seq64, optimize_pgure=0, lambda1=3.0)
    e = G.rel_l2(Yg, Yo)
    log("  rel L2 =", e)
    assert e < FRAME_TOL
    assert np.array_equal(eg, eo)


def test_denoise_pgure_motion_off(gpu):
    # the CPU-reference configuration at reduced size: static blobs, motion off
    log("denoise PGURE motion off")
    clean, noisy = phantom.noisy_sequence(64, 20, drift=0, alpha=0.5, sigma=5.0, seed=1)
    kw = dict(motion_estimation=False)
    Yg, eg = G.denoise(noisy, **kw)
    Yo, eo = O.denoise(noisy, motion_estimation=0)
    # noise estimates come from the same arithmetic: identical
    assert np.array_equal(eg[:, 1:], eo[:, 1:])
    lam_rel = np.abs(eg[:, 0] - eo[:, 0]) / np.abs(eo[:, 0])
    log("  lambda rel diff max =", lam_rel.max())
    # frames at the GPU's lambda against the oracle at the same lambda
    for t in (0, 9, 19):
        g = G.window_reconstruct(noisy, t, eg[t, 0], motion_estimation=False)
        o = O.window_reconstruct(noisy, t, eg[t, 0], motion_estimation=0)
        assert G.rel_l2(g, o) < FRAME_TOL
    assert lam_rel.max() < 1e-2


@pytest.mark.parametrize("mf", [-1, 5])
def test_denoise_pgure_motion_on(gpu, seq64, mf):
    log("denoise PGURE motion on, motion_filter=%d" % mf)
    Yg, eg = G.denoise(seq64, motion_filter=mf)
    Yo, eo = O.denoise(seq64, motion_filter=mf)
    assert np.array_equal(eg[:, 1:], eo[:, 1:])
    lam_rel = np.abs(eg[:, 0] - eo[:, 0]) / np.abs(eo[:, 0])
    log("  lambda rel diff max =", lam_rel.max(), " frames rel L2 =", G.rel_l2(Yg, Yo))
    same = lam_rel == 0
    log("  windows with identical lambda: %d / %d" % (same.sum(), len(same)))
    assert lam_rel.max() < 1e-2
    # frames at the GPU's lambda: first, middle and last window, with their
    # trajectories bit-exact
    nf = seq64.shape[2]
    for t in (0, nf // 2, nf - 1):
        assert np.array_equal(G.window_trajectories(seq64, t, motion_filter=mf),
                              O.window_trajectories(seq64, t, motion_filter=mf))
        g = G.window_reconstruct(seq64, t, eg[t, 0], motion_filter=mf)
        o = O.window_reconstruct(seq64, t, eg[t, 0], motion_filter=mf)
        e = G.rel_l2(g, o)
        log("  t=%d window rel L2 at the GPU lambda = %.3g" % (t, e))
        assert e < FRAME_TOL
        # and the output frame itself is that window's reference frame
        ref = t if t < 7 else (t - (nf - 15) if t >= nf - 7 else 7)
        mx = seq64[:, :, t - ref:t - ref + 15].max()
        assert G.rel_l2(Yg[:, :, t], o[:, :, ref] * mx) < FRAME_TOL


# Every other instantiated (patch, window) family, at a size the oracle
# finishes quickly: fixed-lambda denoise of the whole sequence, and the
# PGURE risk of one window at three thresholds.  Patch 8 with step 2 is the
# reference's non-grid patch subset (src/svt.hpp:61-92).
SHAPES = [
    (2, 1, 3, False),
    (3, 1, 5, True),
    (3, 1, 9, True),
    (4, 2, 15, True),
    (4, 1, 7, False),
    (6, 1, 7, True),
    (6, 2, 15, True),
    (8, 2, 7, True),
    (8, 2, 15, True),
    (8, 1, 31, True),
    (6, 1, 31, True),
]


@pytest.fixture(scope="module")
def seq64x40():
    clean, noisy = phantom.noisy_sequence(64, 40, drift=1, alpha=1.0, sigma=6.0, seed=12)
    return noisy


@pytest.fixture
def svd_mapping():
    """Select an SVD lanes-per-matrix mapping for one test, restore the
    default afterwards."""
    yield G.set_svd_mapping
    G.set_svd_mapping(0)


@pytest.mark.parametrize("mapping", [0, 1])
@pytest.mark.parametrize("bs,step,T,expw", SHAPES)
def test_shapes_fixed_lambda_and_risk(gpu, seq64x40, svd_mapping, bs, step, T, expw, mapping):
    # both lanes-per-matrix mappings of the SVD of each shape (the primary
    # one is the default, the alternative one a choice of the autotuner)
    log("shape", bs, step, T, expw, "mapping", mapping)
    svd_mapping(mapping)
    kw = dict(patch_size=bs, patch_overlap=step, trajectory_length=T, exponential_weighting=expw)
    for t in (0, 20, 39):  # ARPS trajectories bit-exact at this block size
        tg = G.window_trajectories(seq64x40, t, **kw)
        to = O.window_trajectories(seq64x40, t, **kw)
        log("  trajectories t=%d: %d of %d differ" % (t, int((tg != to).any(-1).sum()), tg.shape[0] * tg.shape[1]))
        assert np.array_equal(tg, to)
    lam = 0.5 if expw else 0.05
    Yg, eg = G.denoise(seq64x40, optimize_pgure=False, lambda1=lam, **kw)
    Yo, eo = O.denoise(seq64x40, optimize_pgure=0, lambda1=lam, **kw)
    e = G.rel_l2(Yg, Yo)
    log("  frames rel L2 =", e)
    assert e < FRAME_TOL
    t = 20
    noise = [0.02, 0.03, 0.04]
    rg = G.window_risk(seq64x40, t, noise, [lam / 3, lam, 3 * lam], **kw)
    ro = O.window_risk(seq64x40, t, noise, [lam / 3, lam, 3 * lam], **kw)
    log("  risk rel =", np.abs(rg - ro) / np.abs(ro))
    assert np.all(np.abs(rg - ro) <= RISK_TOL * np.abs(ro))


def _bowl_sequence(n=64, nf=20, seed=5):
    """Every frame but the middle one is a paraboloid bowl: the motion
    search slides each block downhill, so the trajectories pile up (up to
    225 patches on one position of a frame) and tiles of the gather hold
    more patch rows than it stages in LDS (975 against 448); those tiles
    form each contribution where it is summed."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n]
    bowl = (yy - 30.5) ** 2 + (xx - 30.5) ** 2
    X = np.empty((n, n, nf))
    for k in range(nf):
        X[:, :, k] = 1000.0 + (0.0 if k == nf // 2 else bowl) + rng.random((n, n))
    return np.asfortranarray(X)


def test_piled_up_trajectories(gpu):
    X = _bowl_sequence()
    kw = dict(patch_size=4, patch_overlap=1, trajectory_length=15, motion_window=7, motion_filter=-1)
    t = 10
    tg = G.window_trajectories(X, t, **kw)
    to = O.window_trajectories(X, t, **kw)
    assert np.array_equal(tg, to)
    n1 = X.shape[0] - 3
    pile = max(int(np.bincount(tg[k][:, 0] + n1 * tg[k][:, 1]).max()) for k in range(tg.shape[0]))
    log("most patches on one position:", pile)
    assert pile > 100
    for expw, lam in ((True, 0.5), (False, 0.05)):
        noise = [0.02, 0.03, 0.04]
        rg = G.window_risk(X, t, noise, [lam / 3, lam, 3 * lam], exponential_weighting=expw, **kw)
        ro = O.window_risk(X, t, noise, [lam / 3, lam, 3 * lam], exponential_weighting=expw, **kw)
        log("  expw", expw, "risk rel =", np.abs(rg - ro) / np.abs(ro))
        assert np.all(np.abs(rg - ro) <= RISK_TOL * np.abs(ro))
        Yg, _ = G.denoise(X, optimize_pgure=False, lambda1=lam, exponential_weighting=expw, **kw)
        Yo, _ = O.denoise(X, optimize_pgure=0, lambda1=lam, exponential_weighting=expw, **kw)
        e = G.rel_l2(Yg, Yo)
        log("  frames rel L2 =", e)
        assert e < FRAME_TOL


def _profile_json(L):
    n = L.pgsvt_profile_read(None, 0)
    buf = C.create_string_buffer(n)
    assert L.pgsvt_profile_read(buf, n) == 0
    return json.loads(buf.value.decode())


def test_profiling_leaves_results_unchanged(gpu, seq64):
    # the profiler (events around the stages; at level 2 also the count of
    # the ranks each evaluation keeps) must not change a single output value
    L = P.load()
    L.pgsvt_profile_enable.argtypes = [C.c_int]
    L.pgsvt_profile_read.argtypes = [C.c_char_p, C.c_uint64]
    Y0, e0 = G.denoise(seq64)
    try:
        for level in (1, 2):
            log("profile level", level)
            L.pgsvt_profile_enable(level)
            Y, e = G.denoise(seq64)
            prof = _profile_json(L)
            L.pgsvt_profile_enable(0)
            assert np.array_equal(Y, Y0) and np.array_equal(e, e0)
            for stage in ("search_eval", "eval_reconstruct", "gather_risk", "svd_jacobi"):
                assert prof[stage]["launches"] > 0 and prof[stage]["ms"] > 0, stage
            # a round's two stages are timed within the round
            assert prof["eval_reconstruct"]["ms"] + prof["gather_risk"]["ms"] <= prof["search_eval"]["ms"] * (1 + 1e-6)
            assert prof["search_eval"]["launches"] == prof["gather_risk"]["launches"]
            if level == 1:
                assert "eval_kept_ranks" not in prof
            else:
                # every patch-variant of every evaluation keeps between 0 and
                # nw ranks
                kept, pv = prof["eval_kept_ranks"]["units"], prof["eval_reconstruct"]["units"]
                log("  kept ranks %.0f over %.0f patch-variants" % (kept, pv))
                assert 0 < kept <= 15 * pv
    finally:
        L.pgsvt_profile_enable(0)
