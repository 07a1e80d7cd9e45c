"""FP32 against FP64 on the GPU, at the 1024 x 1024, patch 8, step 2 shape.

    python pgure-svt_amd/precision_study.py [--frames 30] [--windows 7,11,15,19,22]

Sequence: the blob phantom at 1024 x 1024 (alpha = 2, sigma = 10, mu = 5),
the first `frames` frames; patch 8, step 2, trajectory 15 (64 x 15 Casorati
matrices, 260,096 patches per window), ARPS on.  For each listed output
frame the library runs the window in FP32 (the production path: FP32 Jacobi,
FP32 factors and reconstructions, FP64 risk sums) and in FP64 (everything
FP64, pgsvt_set_precision(64)), and reports
  * singular values: max over patches of max_r |s32_r - s64_r| / s64_max;
  * reconstruction at the FP64 search's lambda: relative L2 error of the
    FP32 window volume;
  * PGURE risk at that lambda and at lambda / 2, 2 lambda: relative error;
  * the PGURE search itself: the two lambdas and the relative L2 error of
    the two output frames.
Prints one JSON line.
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from pguresvt import _pguresvt as P  # noqa: E402
from pguresvt import learned as Lr  # noqa: E402
from pguresvt import phantom  # noqa: E402

PARAMS = dict(trajectory_length=15, patch_size=8, patch_overlap=2, motion_window=7, motion_filter=-1, noise_method=4,
              max_iter=500, n_jobs=-1, random_seed=4, optimize_pgure=1, exponential_weighting=1,
              motion_estimation=1, lambda1=-1.0, noise_alpha=-1.0, noise_mu=-1.0, noise_sigma=-1.0, tol=1e-7)


def set_precision(bits):
    L = P.load()
    L.pgsvt_set_precision.argtypes = [C.c_int]
    P.check(L.pgsvt_set_precision(bits))


def singular_values(X, t):
    L = P.load()
    X = np.asfortranarray(X)
    nr, nc, nf = X.shape
    p = P.make_params(**PARAMS)
    pall = (nr - 7) * (nc - 7)
    out = np.zeros((pall, 15))
    n = C.c_uint32()
    P.check(L.pgsvt_window_singular_values(X.ctypes.data, 2, nr, nc, nf, C.byref(p), t, out.ctypes.data, C.byref(n)))
    return out[:n.value]


def reconstruct(X, t, lam):
    L = P.load()
    X = np.asfortranarray(X)
    nr, nc, nf = X.shape
    p = P.make_params(**PARAMS)
    vol = np.zeros((nr, nc, 15), order="F")
    P.check(L.pgsvt_window_reconstruct(X.ctypes.data, 2, nr, nc, nf, C.byref(p), t, C.c_double(lam), vol.ctypes.data))
    return vol


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--frames", type=int, default=30)
    ap.add_argument("--windows", default="7,11,15,19,22")
    a = ap.parse_args()
    t0 = time.time()
    _, X = phantom.noisy_sequence(a.n, a.frames, blobs=8, drift=2, alpha=2.0, mu=5.0, sigma=10.0, seed=4)
    rows = []
    for t in [int(v) for v in a.windows.split(",")]:
        r = {"t": t}
        res = {}
        for bits in (64, 32):
            set_precision(bits)
            Y, est = Lr.denoise_range(X, t, t + 1, **PARAMS)
            res[bits] = dict(Y=Y[:, :, 0], est=est[0], s=singular_values(X, t))
        lam = float(res[64]["est"][0])
        lams = [lam / 2, lam, 2 * lam]
        for bits in (64, 32):
            set_precision(bits)
            res[bits]["vol"] = reconstruct(X, t, lam)
            res[bits]["risk"] = Lr.window_risk(X, t, res[64]["est"][1:], lams, **PARAMS)
        s32, s64 = res[32]["s"], res[64]["s"]
        smax = np.maximum(s64[:, :1], 1e-300)
        r["sv_rel_err_max"] = float((np.abs(s32 - s64) / smax).max())
        r["sv_rel_err_median"] = float(np.median((np.abs(s32 - s64) / smax).max(1)))
        r["recon_rel_l2"] = rel_l2(res[32]["vol"], res[64]["vol"])
        r["risk_rel_err"] = [float(abs(x - y) / abs(y)) for x, y in zip(res[32]["risk"], res[64]["risk"])]
        r["lambda_fp32"], r["lambda_fp64"] = float(res[32]["est"][0]), lam
        r["frame_rel_l2_pgure"] = rel_l2(res[32]["Y"], res[64]["Y"])
        r["noise_identical"] = bool(np.array_equal(res[32]["est"][1:], res[64]["est"][1:]))
        rows.append(r)
        print("window %d done, %.0f s" % (t, time.time() - t0), flush=True)
    set_precision(32)
    summary = {
        "shape": [a.n, a.n, a.frames], "patch": 8, "step": 2, "casorati": [64, 15],
        "patches_per_window": int(res[32]["s"].shape[0]),
        "sv_rel_err_max": max(r["sv_rel_err_max"] for r in rows),
        "recon_rel_l2_max": max(r["recon_rel_l2"] for r in rows),
        "risk_rel_err_max": max(max(r["risk_rel_err"]) for r in rows),
        "frame_rel_l2_pgure_max": max(r["frame_rel_l2_pgure"] for r in rows),
        "lambda_rel_diff_max": max(abs(r["lambda_fp32"] / r["lambda_fp64"] - 1) for r in rows),
        "windows": rows, "seconds": time.time() - t0,
    }
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
