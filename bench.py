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

P_SEL = (N - PATCH + 1) ** 2  # patches per window (step 1)

# Algorithmic work per unit of each timed stage (units as the library counts
# them in its profile):
#   svd_jacobi          matrix            thin SVD of one m x n Casorati matrix with V and U S:
#                                         14 m n^2 + 8 n^3 flops (Golub-Van Loan count)
#   divergence_factors  matrix            every datum moved once (fp32): all V^T and W^T rows
#                                         read (4 n (n + m)), the divergence factors written
#                                         (4 n), frame positions (4 n / 3) and delta2/W of the
#                                         window's voxels (4 NVOX / (3 P)) shared by the three
#                                         variants of a patch
#   eval_reconstruct    patch x variant   every datum moved once (fp32 factors and rows): the U
#                                         reconstruction of every frame written (4 m n / 3: one
#                                         of three variants), the singular values (4 n), bucket
#                                         places (4 n / 3); plus, per kept rank (counted by the
#                                         library over the three variants, which keep the same
#                                         ranks but for a few patches), its divergence factor
#                                         (4) and, for the U variant, its V^T and W^T rows
#                                         (4 (n + m) / 3 per counted rank)
#   gather_risk         voxel             every patch row of the U reconstruction read once (fp32),
#                                         U (fp64) and 1/W (fp64) of the voxel read once
EVAL_FIXED = 4.0 * M * NW / 3 + 4.0 * NW + 4.0 * NW / 3
EVAL_PER_RANK = 4.0 * (NW + M) / 3 + 4.0
ALGO = {
    "svd_jacobi": ("mfma", 14 * M * NW**2 + 8 * NW**3),
    "divergence_factors": ("hbm", 4.0 * NW * (NW + M) + 4.0 * NW + 4.0 * NW / 3 + 4.0 * NVOX / (3 * P_SEL)),
    "eval_reconstruct": ("hbm", EVAL_FIXED),
    "gather_risk": ("hbm", 4.0 * P_SEL * M * NW / NVOX + 16),
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


def run_gpu(noisy, steps, warmup, model=None):
    """Time `steps` calls of the device entry point (PGURE search, or the
    learned lambda when a model is given) after `warmup` calls."""
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
    if model is None:
        def step():
            P.check(L.pgsvt_denoise_device(xd, 2, N, N, FRAMES, C.byref(p), yd, est.ctypes.data))
    else:
        from pguresvt import learned as Lr

        L.pgsvt_denoise_learned_device.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                                   C.POINTER(P.pgsvt_params), C.POINTER(Lr._Model), C.c_void_p,
                                                   C.c_void_p]
        m = model._c()

        def step():
            P.check(L.pgsvt_denoise_learned_device(xd, 2, N, N, FRAMES, C.byref(p), C.byref(m), yd, est.ctypes.data))

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


def learned_eval(noisy, steps, warmup, dt_plain, est_plain):
    """The learned-lambda path on the same workload: frames/s, and the PGURE
    risk gap of each window's proposed lambda against the searched one
    (both risks evaluated by the library's risk of that window with the
    search's noise estimates).  Windows whose gap exceeds 0.5 % fall back to
    the search: counted at the search's per-window time."""
    from pguresvt import learned as Lr

    path = Lr.default_model_path()
    if not os.path.exists(path):
        return None
    model = Lr.LambdaModel.load(path)
    dt, prof, _, est = run_gpu(noisy, steps, warmup, model)
    gaps = []
    for t in range(FRAMES):
        r = Lr.window_risk(noisy, t, est_plain[t, 1:], [est_plain[t, 0], est[t, 0]], **PARAMS)
        gaps.append((r[1] - r[0]) / abs(r[0]))
    gaps = np.array(gaps)
    fb = int((gaps > 0.005).sum())
    per_step_plain = dt_plain
    per_step_learned = dt / steps
    with_fb = per_step_learned + fb / FRAMES * per_step_plain
    return {"value": FRAMES / per_step_learned, "unit": "frames/s", "speedup": per_step_plain / per_step_learned,
            "risk_gap_mean": float(gaps.mean()), "risk_gap_median": float(np.median(gaps)),
            "risk_gap_max": float(gaps.max()), "windows": FRAMES, "windows_gap_over_0.5pct": fb,
            "risk_gap_mean_after_fallback": float(np.where(gaps > 0.005, 0.0, gaps).mean()),
            "speedup_with_fallback": per_step_plain / with_fb,
            "lambda_rel_error_median": float(np.median(np.abs(est[:, 0] / est_plain[:, 0] - 1))),
            "kernels_ms_per_step": {k: v["ms"] / steps for k, v in prof.items()}}


def roofline(prof):
    """Roofline of the stage with the largest GPU time in the timed steps."""
    known = {k: v for k, v in prof.items() if k in ALGO and v["launches"] > 0}
    if not known:
        return None, None
    name = max(known, key=lambda k: known[k]["ms"])
    e = known[name]
    bound, per_unit = ALGO[name]
    avg_s = e["ms"] / e["launches"] / 1e3
    units_per_launch = e["units"] / e["launches"]
    work = per_unit * units_per_launch
    kept = prof.get("eval_kept_ranks")
    if name == "eval_reconstruct" and kept:
        work += EVAL_PER_RANK * kept["units"] / e["launches"]
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            t = json.load(f).get("stages", {}).get(name)
        if t and "hbm_bytes_per_unit" in t:
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
    r["units_per_launch"] = units_per_launch
    # Launch sizes vary from run to run (windows join the lambda search as
    # their host noise estimates finish); the time per unit does not.
    r["ps_per_unit"] = e["ms"] * 1e9 / e["units"]
    return r, name


# kernels of each profiled stage, as rocprofv3 names them
STAGE_KERNELS = {
    "svd_jacobi": "pgsvt::k_svd<",
    "divergence_factors": "pgsvt::k_divq<",
    "eval_reconstruct": "pgsvt::k_eval",
    "gather_risk": "pgsvt::k_gather_",
}


def copy_probe(mib):
    """One elementwise copy (read + write of `mib` MiB) for calibrating the
    FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": 4 * n}))


def _pmc_rows(d, counter):
    import csv
    import glob

    rows = []
    for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(fn) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] == counter:
                    rows.append((r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"])))
    if not rows:
        raise SystemExit("no %s rows under %s" % (counter, d))
    return rows


def summarize_pmc(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_bytes, units_json, out):
    """HBM traffic per unit of each stage from FETCH_SIZE / WRITE_SIZE runs
    (KiB per dispatch), each scaled by what the counter reads for a copy of
    `copy_bytes` bytes read and written.  units_json: the bench line of the
    profiled run (for the units each stage processed)."""
    def copy_kib(d, c):
        rows = [r for r in _pmc_rows(d, c) if "elementwise" in r[0]]
        return max(r[2] for r in rows)
    f_fetch = copy_bytes / (1024.0 * copy_kib(copy_fetch_dir, "FETCH_SIZE"))
    f_write = copy_bytes / (1024.0 * copy_kib(copy_write_dir, "WRITE_SIZE"))
    with open(units_json) as f:
        line = json.loads(f.read().strip().splitlines()[-1])
    units = line["stage_units_per_step"]
    steps = line["steps"] + line["warmup"]
    fetch = _pmc_rows(fetch_dir, "FETCH_SIZE")
    write = _pmc_rows(write_dir, "WRITE_SIZE")
    res = {"copy_probe_bytes": copy_bytes, "fetch_scale": f_fetch, "write_scale": f_write, "stages": {}}
    for st, kname in STAGE_KERNELS.items():
        fk = sum(r[2] for r in fetch if kname in r[0])
        wk = sum(r[2] for r in write if kname in r[0])
        nd = sum(1 for r in fetch if kname in r[0])
        if nd == 0 or st not in units:
            continue
        u = units[st] * steps
        hbm = 1024.0 * (fk * f_fetch + wk * f_write)
        res["stages"][st] = {"dispatches": nd, "fetch_kib": fk, "write_kib": wk, "units": u,
                             "hbm_bytes_per_unit": hbm / u,
                             "algorithmic_bytes_per_unit": ALGO[st][1] if ALGO[st][0] == "hbm" else None}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")
    print(json.dumps(res))


def cpu_baseline(noisy, windows):
    """The oracle on `windows` output frames in the middle of the sequence,
    one window per thread, SVDs through LAPACK dgesdd (single-threaded, so
    `cores` counts every thread used)."""
    for v in ("MKL_NUM_THREADS", "OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS"):
        os.environ[v] = "1"
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
                    help="output frames in the CPU-baseline sample, one thread each (default 2: about 30 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-lambda comparison")
    ap.add_argument("--copy-probe", type=int, metavar="MIB", default=None,
                    help="only run one MIB-MiB elementwise copy (counter calibration) and exit")
    ap.add_argument("--summarize-pmc", nargs=6, default=None,
                    metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR", "BENCH_JSON", "OUT"),
                    help="write per-stage HBM traffic from rocprofv3 --pmc csv output and exit")
    a = ap.parse_args()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.summarize_pmc:
        fd, wd, cfd, cwd, bj, out = a.summarize_pmc
        summarize_pmc(fd, wd, cfd, cwd, (a.copy_probe or 1024) * (1 << 20), bj, out)
        return
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the path shards by output frame across GPUs with no exchange, "
              "but this build measures a single MI355X (got --gpus %d)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    clean, noisy = workload()
    dt, prof, Y, est = run_gpu(noisy, a.steps, a.warmup)
    fps = FRAMES * a.steps / dt
    roof, _ = roofline(prof)
    learned = None if a.no_learned else learned_eval(noisy, a.steps, a.warmup, dt / a.steps, est)
    cpu = None
    if not a.no_cpu_baseline:
        cw = a.cpu_windows or max(1, min(2, os.cpu_count() or 1))
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
        "learned_lambda": learned,
        "kernels_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items()},
        "stage_units_per_step": {k: v["units"] / a.steps for k, v in prof.items()},
        "lambda_median": float(np.median(est[:, 0])),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
