"""Shapes beyond the instantiated table, through the C ABI against the CPU
FP64 oracle: patch sizes 5, 7 and 16, windows of any odd length up to 31
frames (the kernels' run-time window form), search windows above 15 pixels,
and the reference's own example run (examples/param_example.svt on
examples/example.tif, fixture tests/golden/example.tif) through the
command-line program."""

import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as O
import pgsvt_lib as G
from pguresvt import phantom
from test_frontend import CLI, tiff_read

pytestmark = pytest.mark.gpu

FRAME_TOL = 1e-5
RISK_TOL = 1e-4
HERE = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def seq48():
    return phantom.noisy_sequence(48, 34, drift=1, alpha=1.0, sigma=6.0, seed=41)[1]


# (patch, step, trajectory, exponential weighting): windows the table has no
# kernels for (patches 2-4 at 1 and 7 frames included), patches 5, 7 and 16,
# steps 1-3, and even clamped trajectory lengths Nt (trajectory_length itself
# must be odd), whose window has Nt + 1 frames (src/pguresvt.hpp:57-58):
# patch 3 clamps 21 to 8 (window 9), patch 5 clamps 27 and 29 to 24 (window
# 25), patch 2 clamps 17 to 3.  With an even Nt the motion estimator runs
# over the window's Nt + 1 frames (src/arps.hpp:35) while the driver places
# the output frame with Nt (src/pguresvt.hpp:161): the end windows (t = 33)
# check both.
SHAPES = [
    (5, 1, 7, True),
    (5, 2, 25, True),
    (7, 1, 11, False),
    (7, 2, 31, True),
    (6, 1, 21, True),
    (8, 2, 19, True),
    (4, 1, 1, True),
    (3, 1, 7, True),
    (2, 1, 1, False),
    (16, 2, 15, True),
    (16, 3, 5, False),
    (3, 2, 21, False),
    (5, 1, 27, True),
    (5, 2, 29, False),
    (2, 1, 17, True),
]


@pytest.mark.parametrize("bs,step,T,expw", SHAPES)
def test_generic_shapes(gpu, seq48, bs, step, T, expw):
    kw = dict(patch_size=bs, patch_overlap=step, trajectory_length=T, exponential_weighting=expw, motion_filter=-1)
    log("shape", bs, step, T, expw)
    for t in (0, 17, 33):
        assert np.array_equal(G.window_trajectories(seq48, t, **kw), O.window_trajectories(seq48, t, **kw))
    lam = 0.5 if expw else 0.05
    Yg, eg = G.denoise(seq48, optimize_pgure=False, lambda1=lam, **kw)
    Yo, eo = O.denoise(seq48, optimize_pgure=0, lambda1=lam, **kw)
    e = G.rel_l2(Yg, Yo)
    log("  frames rel L2 =", e)
    assert e < FRAME_TOL
    noise = [0.02, 0.03, 0.04]
    rg = G.window_risk(seq48, 17, noise, [lam / 3, lam, 3 * lam], **kw)
    ro = O.window_risk(seq48, 17, noise, [lam / 3, lam, 3 * lam], **kw)
    log("  risk rel =", np.abs(rg - ro) / np.abs(ro))
    assert np.all(np.abs(rg - ro) <= RISK_TOL * np.abs(ro))


@pytest.mark.parametrize("mw", [17, 21, 31])
def test_wide_motion_window(gpu, mw):
    # search windows above 15 pixels (up to 31) run the wide ARPS kernel
    _, X = phantom.noisy_sequence(64, 17, drift=4, alpha=1.0, sigma=6.0, seed=mw)
    for t in (0, 8, 16):
        g = G.window_trajectories(X, t, motion_window=mw, motion_filter=-1)
        o = O.window_trajectories(X, t, motion_window=mw, motion_filter=-1)
        assert np.array_equal(g, o)


def test_reference_example_parameters(gpu, tmp_path):
    # examples/param_example.svt (src/PGURE-SVT.cpp:48-99 reads it): frames
    # 1-25 of examples/example.tif, patch 16, overlap 2, trajectory 15,
    # noise 0.1/0.1/0.1, fixed lambda 0.15, ARPS on (window 7, median 5),
    # exponential weighting, seed 1.  The program's output equals the
    # library's frames truncated to uint16, and those equal the oracle's.
    import shutil

    src = str(tmp_path / "example.tif")
    shutil.copy(os.path.join(HERE, "golden", "example.tif"), src)
    params = ("filename : ./example.tif\nstart_frame : 1\nend_frame : 25\npatch_size : 16\npatch_overlap : 2\n"
              "trajectory_length : 15\nnormalize : false\nnoise_alpha : 0.1\nnoise_mu : 0.1\nnoise_sigma : 0.1\n"
              "optimize_pgure : false\nlambda : 0.15\nmotion_estimation : true\nexponential_weighting : true\n"
              "random_seed : 1\n")
    (tmp_path / "param_example.svt").write_text(params)
    r = subprocess.run([CLI, str(tmp_path / "param_example.svt")], capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out, bits = tiff_read(str(tmp_path / "example-CLEANED.tif"))
    X, _ = tiff_read(src)
    X = np.asfortranarray(X[:, :, :25])
    assert bits == 16 and out.shape == X.shape
    kw = dict(patch_size=16, patch_overlap=2, trajectory_length=15, noise_alpha=0.1, noise_mu=0.1, noise_sigma=0.1,
              optimize_pgure=False, lambda1=0.15, motion_estimation=True, motion_window=7, motion_filter=5,
              exponential_weighting=True, random_seed=1, max_iter=1000)
    Yg, eg = G.denoise(X, **kw)
    expect = np.where(Yg > 0, np.minimum(Yg, 65535), 0).astype(np.uint16)
    assert np.array_equal(out, expect)
    kwo = dict(kw, optimize_pgure=0, motion_estimation=1, exponential_weighting=1)
    try:
        O.set_svd(1)
    except RuntimeError:
        pass
    try:
        Yo, eo = O.denoise(X, **kwo)
    finally:
        O.set_svd(0)
    e = G.rel_l2(Yg, Yo)
    log("example: frames rel L2 against the oracle %.3g" % e)
    assert e < FRAME_TOL
