"""Benchmark of the MI355X PGURE-SVT path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-windows C] [--batch-mix S]

Workload: 100 frames of 256x256 fp32 (blob phantom, every blob drifting by
+-2 px per frame, Poisson-Gaussian noise alpha=1, sigma=8), ARPS motion
estimation on (4x4 blocks = the patch, search window 7), patch 4, step 1,
trajectory 15 (16x15 Casorati matrices), no median pre-filter, lambda from
the PGURE search per frame with noise parameters estimated.

One step = one pgsvt_denoise_device call over the whole sequence with the
input already resident in HBM and the output written to HBM (`value`).  The
same call with host buffers (pgsvt_denoise: the H2D copy of the input and the
D2H copy of the output included) is timed separately and reported as
`host_buffers` (median of its steps).  Prints one JSON line.  The roofline
object is for the kernel with the largest share of GPU time in the timed
steps, measured with HIP events on the library's stream, with the
algorithmic work per unit of the survey's measurement plan: 14 m n^2 + 8 n^3
flops per m x n SVD (FP32 vector work), 40 bytes per voxel per lambda
evaluation of a window.  The CPU baseline is the CPU FP64 oracle (a
restatement of the reference CPU path using LAPACK dgesdd for the SVDs) on a
sample of output frames, one window per host core.  `batch_mix` times the
batched uneven workload (independent sequences of mixed shapes, lambda by a
64-candidate PGURE grid) on its first S sequences.
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

# Algorithmic work per unit used for the roofline (units as the library
# counts them in its profile), from the survey's measurement plan:
#   svd_jacobi          matrix  14 m n^2 + 8 n^3 flops (Golub-Van Loan thin SVD), FP32 vector work
#   lambda evaluation   voxel   40 bytes: the four reconstructed volumes written and read back
#                               (2 * 4 * 4) plus U and W read (8), per voxel of the window
# (the library's lambda evaluation materialises none of the four volumes: the
# bytes are the algorithm's, not what the kernels move).
SURVEY_ALGO = {
    "svd_jacobi": ("valu", 14 * M * NW**2 + 8 * NW**3),
    "gather_risk": ("hbm", 40.0),
    "search_eval": ("hbm", 40.0),
}

# The design's own per-unit data movement of each stage (every datum the
# kernels read or write moved once), kept for the traffic comparison:
#   svd_jacobi          matrix            thin SVD of one m x n Casorati matrix with V and U S:
#                                         14 m n^2 + 8 n^3 flops (Golub-Van Loan count)
#   divergence_factors  matrix            every datum moved once (fp32): all V^T and W^T rows
#                                         read (4 n (n + m)), the divergence factors written
#                                         (4 n), frame positions (4 n / 3) and delta2/W of the
#                                         window's voxels (4 NVOX / (3 P)) shared by the three
#                                         variants of a patch
#   eval_reconstruct    patch x variant   every datum moved once (fp32): the first singular value
#                                         of each variant and the one that ends the kept prefix
#                                         (8); plus per kept rank its singular value and divergence
#                                         factor (8) and, for the U variant's further ranks, d_r
#                                         written (4 / 3 per counted rank)
#   static_rows         bucket place      per (patch, frame): v_0[k] and the place read (8), the
#                                         patch's w_0 row and (s_0, s_1) read once per patch
#                                         ((4 m + 8) / n), the row w_0 v_0[k] and the pair
#                                         written (4 m + 8)
#   gather_risk         voxel             every staged row read once (fp32): per (patch, frame)
#                                         its static row w_0 v_0[k] (4 m), its round record (8)
#                                         and patch index (4), P n / NVOX of them per voxel; U
#                                         (fp64) and 1/W (fp64) of the voxel read once
EVAL_FIXED = 8.0
ALGO = {
    "svd_jacobi": ("valu", 14 * M * NW**2 + 8 * NW**3),
    "divergence_factors": ("hbm", 4.0 * NW * (NW + M) + 4.0 * NW + 4.0 * NW / 3 + 4.0 * NVOX / (3 * P_SEL)),
    "eval_reconstruct": ("hbm", EVAL_FIXED),
    "static_rows": ("hbm", 8.0 + (4.0 * M + 8.0) / NW + 4.0 * M + 8.0),
    "gather_risk": ("hbm", (4.0 * M + 12.0) * P_SEL * NW / NVOX + 16),
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


def run_host(noisy, steps, warmup):
    """Per-step wall times of pgsvt_denoise with host buffers (input H2D and
    output D2H inside the call)."""
    from pguresvt import _pguresvt as P

    L = P.load()
    X = np.asfortranarray(noisy.astype(np.float32))
    p = P.make_params(**PARAMS)
    Y = np.zeros(X.shape, order="F")
    est = np.zeros((FRAMES, 4), order="F")
    times = []
    for k in range(warmup + steps):
        t0 = time.perf_counter()
        P.check(L.pgsvt_denoise(X.ctypes.data, 2, N, N, FRAMES, C.byref(p), Y.ctypes.data, est.ctypes.data))
        if k >= warmup:
            times.append(time.perf_counter() - t0)
    return times


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
    step_times = []
    for _ in range(steps):
        ts = time.perf_counter()
        step()  # returns after the output is complete on the device
        step_times.append(time.perf_counter() - ts)
    hip.sync()
    dt = time.perf_counter() - t0
    run_gpu.step_times = step_times
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


LEARNED_SEEDS = (900000, 900001, 900002)


def learned_eval(steps, warmup, seeds=LEARNED_SEEDS):
    """The learned-lambda path against the plain search on held-out
    sequences of the bench workload's shape (phantom seeds outside the
    training range): per sequence, both timed over `steps` calls after
    `warmup`; the learned path with the library's own probe and fallback
    (windows whose probe predicts a relative risk gap above the limit are
    searched inside the call, and their time is in the call's).  The risk
    gap of each window: PGURE risk of the learned path's final lambda
    against the searched lambda's, both with the search's noise estimates
    (the library's window risk)."""
    from pguresvt import learned as Lr
    from pguresvt import phantom

    path = Lr.default_model_path()
    if not os.path.exists(path):
        return None
    model = Lr.LambdaModel.load(path)
    Lr.set_fallback()
    per_seed = []
    t_plain = t_learned = 0.0
    all_gaps = []
    for seed in seeds:
        _, noisy = phantom.noisy_sequence(N, FRAMES, blobs=8, drift=2, alpha=1.0, mu=0.0, sigma=8.0, seed=seed)
        dp, _, _, est_p = run_gpu(noisy, steps, warmup)
        dl, prof, _, est = run_gpu(noisy, steps, warmup, model)
        fb = Lr.last_fallbacks()
        gaps = []
        for t in range(FRAMES):
            r = Lr.window_risk(noisy, t, est_p[t, 1:], [est_p[t, 0], est[t, 0]], **PARAMS)
            gaps.append((r[1] - r[0]) / abs(r[0]))
        gaps = np.array(gaps)
        all_gaps.append(gaps)
        t_plain += dp
        t_learned += dl
        per_seed.append({"seed": seed, "plain_frames_per_s": FRAMES * steps / dp,
                         "learned_frames_per_s": FRAMES * steps / dl, "speedup": dp / dl,
                         "fallback_windows": fb, "risk_gap_mean": float(gaps.mean()),
                         "risk_gap_max": float(gaps.max()),
                         "windows_gap_over_0.5pct": int((gaps > 0.005).sum()),
                         "kernels_ms_per_step": {k: v["ms"] / steps for k, v in prof.items()}})
    g = np.concatenate(all_gaps)
    return {"value": FRAMES * steps * len(seeds) / t_learned, "unit": "frames/s",
            "speedup_with_fallback": t_plain / t_learned, "risk_gap_mean": float(g.mean()),
            "risk_gap_median": float(np.median(g)), "risk_gap_max": float(g.max()),
            "windows": int(g.size), "windows_gap_over_0.5pct": int((g > 0.005).sum()),
            "fallback_windows": int(sum(s["fallback_windows"] for s in per_seed)),
            "probe": {"delta": Lr.FALLBACK_DELTA, "gap_limit": Lr.FALLBACK_GAP, "refine": Lr.FALLBACK_REFINE},
            "seeds": per_seed,
            "note": "held-out 256x256x100 sequences (phantom seeds %s); speedup = summed wall time of the plain "
                    "PGURE search over that of the learned path with its own probe and fallback searches; "
                    "gaps judged by the library's window risk with the search's noise estimates"
                    % (", ".join(str(x) for x in seeds))}


def roofline_of(prof, name):
    """Roofline of one profiled stage from the survey's algorithmic work per
    unit and the stage's average launch time."""
    e = prof[name]
    bound, per_unit = SURVEY_ALGO[name]
    avg_s = e["ms"] / e["launches"] / 1e3
    units_per_launch = e["units"] / e["launches"]
    work = per_unit * units_per_launch
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
        r = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
             "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic}
    r["kernel"] = name
    r["algorithmic_per_unit"] = per_unit
    r["avg_launch_ms"] = e["ms"] / e["launches"]
    r["units_per_launch"] = units_per_launch
    # Launch sizes vary from run to run (windows join the lambda search as
    # their host noise estimates finish); the time per unit does not.
    r["ps_per_unit"] = e["ms"] * 1e9 / e["units"]
    return r


def roofline(prof):
    """Roofline of the stage with the largest GPU time in the timed steps,
    and of every stage with an algorithmic figure."""
    known = {k: v for k, v in prof.items() if k in SURVEY_ALGO and v["launches"] > 0}
    if not known:
        return None, {}
    name = max(known, key=lambda k: known[k]["ms"])
    return roofline_of(prof, name), {k: roofline_of(prof, k) for k in known}


# kernels of each profiled stage, as rocprofv3 names them
STAGE_KERNELS = {
    "svd_jacobi": ("pgsvt::k_svd<",),
    "divergence_factors": ("pgsvt::k_divq<",),
    "static_rows": ("pgsvt::k_static_rows<",),
    "eval_reconstruct": ("pgsvt::k_round_prep<", "pgsvt::k_eval<"),
    "gather_risk": ("pgsvt::k_gather_",),
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
    for st, knames in STAGE_KERNELS.items():
        def of(r):
            return any(k in r[0] for k in knames)
        fk = sum(r[2] for r in fetch if of(r))
        wk = sum(r[2] for r in write if of(r))
        nd = sum(1 for r in fetch if of(r))
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


def host_cores():
    """CPUs this process may run on, capped by a cgroup CPU quota (the
    library's own host thread count)."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max" and float(period) > 0:
            n = min(n, max(1, int(np.ceil(float(quota) / float(period)))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
            "cpu_model": cpu_model(), "host_cores": host_cores(),
            "sample": "output frames %d-%d of the same sequence (%d full windows: noise estimate, ARPS, 4 SVD sets, "
                      "SBPLX lambda search), one window per thread on %d of the host's %d cores, SVD %s, "
                      "%.1f s wall" % (t0, t0 + windows - 1, windows, windows, host_cores(), svd, dt)}, \
        est[t0:t0 + windows, 0]


def batch_mix(n_seq, group=4, n_lambda=64):
    """The batched uneven workload: the first n_seq sequences of the seed-3
    draw (N x N x F, N in {128, 256, 512}, F in {30, 60, 200}, patch 4/6/8,
    trajectory 7/15/31), 64 lambda candidates per window, through
    pgsvt_denoise_batch with `group` sequences in flight.  The sequences are
    made before the timed call; host buffers in and out (the copies, each
    shape's graph capture and SVD-mapping tuning are inside)."""
    import batch_bench as bb
    from pguresvt import batch as B
    jobs = bb.draw(n_seq)
    items = [bb.make(j) for j in jobs]
    t0 = time.perf_counter()
    outs = B.denoise_batch(items, n_lambda=n_lambda, concurrency=group)
    dt = time.perf_counter() - t0
    frames = sum(X.shape[2] for X, _ in items)
    voxels = sum(X.size for X, _ in items)
    for (X, _), (Y, est) in zip(items, outs):
        if not (np.isfinite(Y).all() and (est[:, 0] > 0).all()):
            raise RuntimeError("batch mix: a sequence came back with non-finite frames or no lambda")
    return {"value": frames / dt, "unit": "frames/s", "voxels_per_s": voxels / dt, "seconds": dt,
            "sequences": n_seq, "frames": frames, "voxels": voxels, "lambda_candidates": n_lambda,
            "in_flight": group,
            "jobs": ["%dx%dx%d p%d T%d" % (j["n"], j["n"], j["frames"], j["patch"], j["trajectory"]) for j in jobs],
            "note": "first %d sequences of the seed-3 draw of the batched uneven workload (32 in the full draw), "
                    "lambda by the least PGURE risk over %d candidates per window, one captured HIP graph per "
                    "batch of windows; host buffers, one timed pgsvt_denoise_batch call" % (n_seq, n_lambda)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-windows", type=int, default=None,
                    help="output frames in the CPU-baseline sample, one thread each (default: one per host core)")
    ap.add_argument("--host-steps", type=int, default=5,
                    help="timed calls with host buffers (H2D and D2H included), median reported")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-lambda comparison")
    ap.add_argument("--learned-only", type=int, metavar="SEED", default=None,
                    help="time only the learned-lambda path (with its probe and fallback) on the held-out "
                         "sequence of that phantom seed, and print its stage times (a profiling target)")
    ap.add_argument("--batch-mix", type=int, default=8, metavar="S",
                    help="sequences of the batched uneven workload timed for the batch_mix field (0: skip)")
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
    if a.learned_only is not None:
        from pguresvt import learned as Lr
        from pguresvt import phantom

        model = Lr.LambdaModel.load(Lr.default_model_path())
        Lr.set_fallback()
        _, noisy = phantom.noisy_sequence(N, FRAMES, blobs=8, drift=2, alpha=1.0, mu=0.0, sigma=8.0,
                                          seed=a.learned_only)
        dt, prof, _, _ = run_gpu(noisy, a.steps, a.warmup, model)
        print(json.dumps({"metric": "learned-lambda frames/s", "value": FRAMES * a.steps / dt, "unit": "frames/s",
                          "seed": a.learned_only, "steps": a.steps, "warmup": a.warmup,
                          "fallback_windows": Lr.last_fallbacks(),
                          "kernels_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items()}}))
        return
    clean, noisy = workload()
    dt, prof, Y, est = run_gpu(noisy, a.steps, a.warmup)
    step_times = run_gpu.step_times
    fps = FRAMES * a.steps / dt
    roof, stages = roofline(prof)
    host = None
    if a.host_steps > 0:
        ht = run_host(noisy, a.host_steps, 1)
        host = {"value": FRAMES / float(np.median(ht)), "unit": "frames/s", "ms_median": 1e3 * float(np.median(ht)),
                "ms_all": [1e3 * t for t in ht],
                "note": "pgsvt_denoise with host buffers: the H2D copy of the input and the D2H copy of the output "
                        "are inside each timed call; median of the calls after one warm-up call"}
    learned = None if a.no_learned else learned_eval(a.steps, a.warmup)
    mix = batch_mix(a.batch_mix) if a.batch_mix > 0 else None
    cpu = None
    if not a.no_cpu_baseline:
        cw = a.cpu_windows or host_cores()
        cpu, _ = cpu_baseline(noisy, cw)
    out = {
        "metric": METRIC,
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "ms_per_step_median": 1e3 * float(np.median(step_times)),
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
        "roofline_stages": stages,
        "host_buffers": host,
        "cpu_baseline": cpu,
        "learned_lambda": learned,
        "batch_mix": mix,
        "kernels_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items()},
        "stage_units_per_step": {k: v["units"] / a.steps for k, v in prof.items()},
        "lambda_median": float(np.median(est[:, 0])),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
