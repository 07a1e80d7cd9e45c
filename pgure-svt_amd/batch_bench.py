"""Batched uneven workload: 32 independent sequences through the lambda-grid
path, each batch of windows one captured HIP graph.

    python pgure-svt_amd/batch_bench.py [--sequences 32] [--group 4] [--compare]

Sequences (seed 3): sizes N x N x F with N in {128, 256, 512}, F in {30,
60, 200}, patch p in {4, 6, 8}, trajectory T in {7, 15, 31} drawn uniformly;
a draw whose clamped window is longer than F is drawn again.  Blob phantom
drifting +-2 px per frame, Poisson-Gaussian noise alpha = 1, sigma = 8.
ARPS on, step 1, exponential weighting, noise estimated, 64 lambda
candidates per window.  Jobs go to pgsvt_denoise_batch `group` at a time
(that many sequences in flight, largest first); progress after each group.
--compare also runs the first group again through the grid (warm: graphs
captured, mappings tuned) and through the reference search
(pgsvt_denoise_batch with n_lambda = 0).  Prints one JSON line.
"""

import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from pguresvt import batch as B  # noqa: E402
from pguresvt import learned as Lr  # noqa: E402
from pguresvt import phantom  # noqa: E402


def draw(n_seq, seed=3):
    rng = np.random.default_rng(seed)
    jobs = []
    while len(jobs) < n_seq:
        n = int(rng.choice([128, 256, 512]))
        f = int(rng.choice([30, 60, 200]))
        p = int(rng.choice([4, 6, 8]))
        t = int(rng.choice([7, 15, 31]))
        if Lr.window_length(p, t) > f:
            continue
        jobs.append(dict(n=n, frames=f, patch=p, trajectory=t, seed=1000 + len(jobs)))
    return jobs


def make(job):
    _, X = phantom.noisy_sequence(job["n"], job["frames"], blobs=8, drift=2, alpha=1.0, mu=0.0, sigma=8.0,
                                  seed=job["seed"])
    params = dict(Lr.TRAIN_PARAMS, patch_size=job["patch"], trajectory_length=job["trajectory"], random_seed=3)
    return X, params


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sequences", type=int, default=32)
    ap.add_argument("--group", type=int, default=4)
    ap.add_argument("--lambdas", type=int, default=64)
    ap.add_argument("--compare", action="store_true")
    a = ap.parse_args()
    jobs = draw(a.sequences)
    t_lib, frames, voxels = 0.0, 0, 0
    t_ref = None
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=a.group) as ex:
        for g0 in range(0, len(jobs), a.group):
            grp = jobs[g0:g0 + a.group]
            items = list(ex.map(make, grp))
            t1 = time.perf_counter()
            outs = B.denoise_batch(items, n_lambda=a.lambdas, concurrency=a.group)
            dt = time.perf_counter() - t1
            t_lib += dt
            for (X, _), (Y, est) in zip(items, outs):
                assert np.isfinite(Y).all() and (est[:, 0] > 0).all()
                frames += X.shape[2]
                voxels += X.size
            if a.compare and g0 == 0:
                # the first grid call captures the graphs and tunes the
                # mappings: time it again warm, then the search (warm too)
                t1 = time.perf_counter()
                B.denoise_batch(items, n_lambda=a.lambdas, concurrency=a.group)
                grid_warm = time.perf_counter() - t1
                t1 = time.perf_counter()
                B.denoise_batch(items, n_lambda=0, concurrency=a.group)
                t_ref = time.perf_counter() - t1
                grid_first = dt
            print("group %d: %d sequences, %.2f s in the library, %.0f s total" %
                  (g0 // a.group, len(grp), dt, time.time() - t0), flush=True)
    out = {"sequences": len(jobs), "frames": frames, "voxels": voxels, "library_s": t_lib,
           "frames_per_s": frames / t_lib, "voxels_per_s": voxels / t_lib, "lambdas": a.lambdas,
           "jobs": jobs}
    if t_ref is not None:
        out["first_group_grid_s"] = grid_first
        out["first_group_grid_warm_s"] = grid_warm
        out["first_group_search_s"] = t_ref
    print(json.dumps(out))


if __name__ == "__main__":
    main()
