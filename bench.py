"""Benchmark of the MI355X PGURE-SVT path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-windows C]

Workload: 100 frames of 256x256 fp32 (blob phantom, every blob drifting by
+-2 px per frame, Poisson-Gaussian noise alpha=1, sigma=8), ARPS motion
estimation on (4x4 blocks = the patch, search window 7), patch 4, step 1,
trajectory 15 (16x15 Casorati matrices), no median pre-filter, lambda from
the PGURE search per frame with noise parameters estimated.

One step = one pgsvt_denoise_device call over the whole sequence with the
input already resident in HBM and the output written to HBM.  Prints one JSON
line.  The roofline object is for the kernel with the largest share of GPU
time in the timed steps, measured with HIP events on the library's stream.
The CPU baseline is the CPU FP64 oracle (a restatement of the reference CPU
path using LAPACK dgesdd for the SVDs) on a sample of output frames, one
window per host thread.
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pgure-svt_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "denoised frames/s @256²×100 (PGURE λ) + learned-λ vs plain GPU speedup @≤0.5% risk gap"

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3

N, FRAMES, PATCH, TRAJ = 256, 100, 4, 15
M, NW = PATCH * PATCH, TRAJ
NVOX = N * N * NW

# algorithmic work per unit of each timed kernel (units as the library counts them)
#   svd_jacobi         matrix      Golub-Van Loan thin SVD: 14 m n^2 + 8 n^3 flops
#   recon_risk         patch x variant   threshold + U diag(s') V^T: 2 m n^2 + m n flops
#   pgure_risk         voxel       3 reconstructions (fp32) + U (fp64) + weight (int32) + delta2 sign: 25 bytes
#   arps               block search      ref + candidate block reads are cache-resident; not a roofline kernel
ALGO = {
    "svd_jacobi": ("mfma", 14 * M * NW**2 + 8 * NW**3),
    "recon_risk": ("mfma", 2 * M * NW**2 + M * NW),
    "recon_final_frame": ("mfma", 2 * M * NW + M),
    "pgure_risk": ("hbm", 25),
}


class Hip:
    """Minimal HIP runtime access for device buffers."""

    def __init__(self):
        self.L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed: hip error %d" % (what, rc))

    def malloc(self, n):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), n), "hipMalloc")
        return p.value

    def free(self, p):
        self.L.hipFree(p)

    def h2d(self, dst, arr):
        self.check(self.L.hipMemcpy(dst, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")

    def d2h(self, arr, src):
        self.check(self.L.hipMemcpy(arr.ctypes.data, src, arr.nbytes, 2), "hipMemcpy D2H")

    def sync(self):
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")


def workload():
    from pguresvt import phantom

    clean, noisy = phantom.noisy_sequence(N, FRAMES, blobs=8, drift=2, alpha=1.0, mu=0.0, sigma=8.0, seed=2)
    return clean, noisy


PARAMS = dict(trajectory_length=TRAJ, patch_size=PATCH, patch_overlap=1, motion_window=7, motion_filter=-1,
              noise_method=4, max_iter=500, n_jobs=-1, random_seed=2, optimize_pgure=True,
              exponential_weighting=True, motion_estimation=True, lambda1=-1.0, noise_alpha=-1.0, noise_mu=-1.0,
              noise_sigma=-1.0, tol=1e-7)


def run_gpu(noisy, steps, warmup):
    from pguresvt import _pguresvt as P

    L = P.load()
    L.pgsvt_denoise_device.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                       C.POINTER(P.pgsvt_params), C.c_void_p, C.c_void_p]
    L.pgsvt_profile_enable.argtypes = [C.c_int]
    L.pgsvt_profile_read.argtypes = [C.c_char_p, C.c_uint64]
    hip = Hip()
    X = np.asfortranarray(noisy.astype(np.float32))
    xd = hip.malloc(X.nbytes)
    yd = hip.malloc(8 * X.size)
    hip.h2d(xd, X)
    p = P.make_params(**PARAMS)
    est = np.zeros((FRAMES, 4), order="F")

    def step():
        P.check(L.pgsvt_denoise_device(xd, 2, N, N, FRAMES, C.byref(p), yd, est.ctypes.data))

    for _ in range(warmup):
        step()
    hip.sync()
    L.pgsvt_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    hip.sync()
    dt = time.perf_counter() - t0
    n = L.pgsvt_profile_read(None, 0)
    buf = C.create_string_buffer(n)
    L.pgsvt_profile_read(buf, n)
    L.pgsvt_profile_enable(0)
    prof = json.loads(buf.value.decode())
    Y = np.zeros(X.shape, order="F")
    hip.d2h(Y, yd)
    hip.free(xd)
    hip.free(yd)
    return dt, prof, Y, est.copy()


def roofline(prof):
    known = {k: v for k, v in prof.items() if k in ALGO and v["launches"] > 0}
    if not known:
        return None, None
    name = max(known, key=lambda k: known[k]["ms"])
    e = known[name]
    bound, per_unit = ALGO[name]
    avg_s = e["ms"] / e["launches"] / 1e3
    units_per_launch = e["units"] / e["launches"]
    work = per_unit * units_per_launch
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s.json" % name)
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f)
        if "hbm_bytes_per_unit" in t:
            traffic = t["hbm_bytes_per_unit"] * units_per_launch
    if bound == "hbm":
        achieved = work / avg_s / 1e9
        r = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
             "frac": achieved / PEAK_HBM_GBS, "traffic": traffic}
    else:
        achieved = work / avg_s / 1e12
        r = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
             "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic}
    r["kernel"] = name
    r["avg_launch_ms"] = e["ms"] / e["launches"]
    return r, name


def cpu_baseline(noisy, windows):
    """The oracle on `windows` output frames in the middle of the sequence,
    one window per thread, SVDs through LAPACK dgesdd."""
    import oracle_lib as O

    try:
        svd = O.set_svd(1)
    except RuntimeError:
        svd = O.set_svd(0)
    t0 = FRAMES // 2 - windows // 2
    kw = dict(PARAMS)
    kw["n_jobs"] = windows
    kw["optimize_pgure"] = 1
    kw["exponential_weighting"] = 1
    kw["motion_estimation"] = 1
    st = time.perf_counter()
    Y, est = O.denoise_range(noisy, t0, t0 + windows, **kw)
    dt = time.perf_counter() - st
    O.set_svd(0)
    return {"value": windows / dt, "unit": "frames/s", "cores": windows, "kind": "port",
            "sample": "output frames %d-%d of the same sequence (%d full windows: noise estimate, ARPS, 4 SVD sets, "
                      "SBPLX lambda search), one window per thread, SVD %s, %.1f s wall"
                      % (t0, t0 + windows - 1, windows, svd, dt)}, est[t0:t0 + windows, 0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-windows", type=int, default=None,
                    help="output frames in the CPU-baseline sample (default: host cores, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the path shards by output frame across GPUs with no exchange, "
              "but this build measures a single MI355X (got --gpus %d)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    clean, noisy = workload()
    dt, prof, Y, est = run_gpu(noisy, a.steps, a.warmup)
    fps = FRAMES * a.steps / dt
    roof, _ = roofline(prof)
    cpu = None
    if not a.no_cpu_baseline:
        cw = a.cpu_windows or max(1, min(16, os.cpu_count() or 1))
        cpu, _ = cpu_baseline(noisy, cw)
    out = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "256x256x100 fp32 drifting-blob sequence, ARPS on (window 7), patch 4, step 1, "
                               "trajectory 15, PGURE lambda per frame, noise estimated",
                   "frames": FRAMES, "frame": [N, N], "patch": PATCH, "trajectory": TRAJ, "casorati": [M, NW],
                   "patches_per_window": (N - PATCH + 1) ** 2},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items()},
        "lambda_median": float(np.median(est[:, 0])),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
