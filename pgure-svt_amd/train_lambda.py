"""Train the learned-lambda proposal network on the GPU.

    python pgure-svt_amd/train_lambda.py [--sets 128:20000:10000,256:6000:40000]
        [--epochs 150] [--out pgure-svt_amd/pguresvt/lambda_model.npz] [--report FILE]

Data: seeded synthetic sequences of N x N x 30, one set per `--sets` entry
N:count:first_seed (sequence i of a set has seed first_seed + i; blob
phantom with 4-16 blobs drifting +-2 px per frame, Poisson-Gaussian noise
with alpha in [0.1, 4] and sigma in [1, 20]).  For each sequence the window
of output frame 15 is labelled with the lambda of the GPU PGURE search
(pgsvt_denoise_range over that frame) and described by the learned-path
features (pgsvt_learned_features).  The last 10 % of each set are held out.
The network is trained with the library's own GPU kernels (Adam, BF16
matrix cores).  Prints one JSON report line.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from pguresvt import learned as Lr  # noqa: E402


def parse_sets(spec):
    sets = []
    for item in spec.split(","):
        n, count, first = (int(v) for v in item.split(":"))
        sets.append((n, count, first))
    return sets


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sets", default="128:20000:10000,256:6000:40000",
                    help="comma-separated N:count:first_seed sets of N x N x 30 sequences")
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=Lr.default_model_path())
    ap.add_argument("--report", default=None)
    ap.add_argument("--workers", type=int, default=4, help="sequences labelled concurrently")
    a = ap.parse_args()
    sets = parse_sets(a.sets)

    # Sequences are labelled by `workers` host threads at once (each library
    # call has its own stream; host cores are shared between the calls).
    from concurrent.futures import ThreadPoolExecutor

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    jobs = max(1, min(16, cores) // a.workers)

    def label(job):
        n, seed = job
        X = Lr.training_sequence(seed, n=n)
        return Lr.label_window(X, 15, n_jobs=jobs)

    total = sum(c for _, c, _ in sets)
    F = np.zeros((total, Lr.n_features(15)), dtype=np.float32)
    y = np.zeros(total)
    train_idx, val_idx = [], []
    t0 = time.time()
    row = 0
    with ThreadPoolExecutor(max_workers=a.workers) as ex:
        for n, count, first in sets:
            n_val = max(1, count // 10)
            for i, (f, lam) in enumerate(ex.map(label, [(n, first + i) for i in range(count)])):
                F[row], y[row] = f, lam
                (val_idx if i >= count - n_val else train_idx).append(row)
                row += 1
                if row % 500 == 0:
                    print("labelled %d / %d sequences (%dx%d set), %.0f s" % (row, total, n, n, time.time() - t0),
                          flush=True)
    y = np.log(np.maximum(y, 1e-6))
    t_label = time.time() - t0

    tr, va = np.array(train_idx), np.array(val_idx)
    t1 = time.time()
    model, losses = Lr.fit(F[tr], y[tr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=0)
    t_train = time.time() - t1
    pred = model.predict_log_lambda(F[va])
    base = np.full_like(y[va], y[tr].mean())
    report = {
        "sets": [{"n": n, "sequences": c, "first_seed": s0} for n, c, s0 in sets],
        "sequences": total, "features": int(F.shape[1]), "label_s": t_label, "train_s": t_train,
        "train_loss_first": float(losses[0]), "train_loss_last": float(losses[-1]),
        "val_rmse_log_lambda": float(np.sqrt(np.mean((pred - y[va]) ** 2))),
        "val_rmse_log_lambda_constant": float(np.sqrt(np.mean((base - y[va]) ** 2))),
        "val_median_rel_lambda_error": float(np.median(np.abs(np.exp(pred - y[va]) - 1))),
        "log_lambda_range": [float(y.min()), float(y.max())],
    }
    # held-out error per set
    start = 0
    per_set = []
    for n, count, _ in sets:
        n_val = max(1, count // 10)
        v = np.arange(start + count - n_val, start + count)
        pv = model.predict_log_lambda(F[v])
        per_set.append({"n": n, "val_rmse_log_lambda": float(np.sqrt(np.mean((pv - y[v]) ** 2))),
                        "val_median_rel_lambda_error": float(np.median(np.abs(np.exp(pv - y[v]) - 1)))})
        start += count
    report["per_set"] = per_set
    model.save(a.out)
    line = json.dumps(report)
    print(line)
    if a.report:
        with open(a.report, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
