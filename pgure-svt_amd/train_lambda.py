"""Train the learned-lambda proposal network on the GPU.

    python pgure-svt_amd/train_lambda.py [--sequences 20000] [--epochs 150]
        [--out pgure-svt_amd/pguresvt/lambda_model.npz] [--report FILE]

Data: seeded synthetic sequences of 128 x 128 x 30 (seeds 10000 + i; blob
phantom with 4-16 blobs drifting +-2 px per frame, Poisson-Gaussian noise
with alpha in [0.1, 4] and sigma in [1, 20]).  For each sequence the window
of output frame 15 is labelled with the lambda of the GPU PGURE search
(pgsvt_denoise_range over that frame) and described by the learned-path
features (pgsvt_learned_features).  10 % of the sequences are held out.
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sequences", type=int, default=20000)
    ap.add_argument("--first-seed", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=Lr.default_model_path())
    ap.add_argument("--report", default=None)
    ap.add_argument("--workers", type=int, default=4, help="sequences labelled concurrently")
    a = ap.parse_args()

    # Sequences are labelled by `workers` host threads at once (each library
    # call has its own stream; host cores are shared between the calls).
    from concurrent.futures import ThreadPoolExecutor

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    jobs = max(1, min(16, cores) // a.workers)

    def label(i):
        X = Lr.training_sequence(a.first_seed + i)
        return Lr.label_window(X, 15, n_jobs=jobs)

    t0 = time.time()
    F = np.zeros((a.sequences, Lr.n_features(15)), dtype=np.float32)
    y = np.zeros(a.sequences)
    with ThreadPoolExecutor(max_workers=a.workers) as ex:
        for i, (f, lam) in enumerate(ex.map(label, range(a.sequences))):
            F[i], y[i] = f, lam
            if (i + 1) % 500 == 0:
                print("labelled %d / %d sequences, %.0f s" % (i + 1, a.sequences, time.time() - t0), flush=True)
    y = np.log(np.maximum(y, 1e-6))
    t_label = time.time() - t0

    n_val = max(1, a.sequences // 10)
    tr, va = slice(0, a.sequences - n_val), slice(a.sequences - n_val, a.sequences)
    t1 = time.time()
    model, losses = Lr.fit(F[tr], y[tr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=0)
    t_train = time.time() - t1
    pred = model.predict_log_lambda(F[va])
    base = np.full_like(y[va], y[tr].mean())
    report = {
        "sequences": a.sequences, "features": int(F.shape[1]), "label_s": t_label, "train_s": t_train,
        "train_loss_first": float(losses[0]), "train_loss_last": float(losses[-1]),
        "val_rmse_log_lambda": float(np.sqrt(np.mean((pred - y[va]) ** 2))),
        "val_rmse_log_lambda_constant": float(np.sqrt(np.mean((base - y[va]) ** 2))),
        "val_median_rel_lambda_error": float(np.median(np.abs(np.exp(pred - y[va]) - 1))),
        "log_lambda_range": [float(y.min()), float(y.max())],
    }
    model.save(a.out)
    line = json.dumps(report)
    print(line)
    if a.report:
        with open(a.report, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
