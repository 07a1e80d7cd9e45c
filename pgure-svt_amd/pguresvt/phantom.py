"""Seeded synthetic microscopy sequences for the benchmark and parity workloads.

Clean signal: a background of 10 plus K Gaussian blobs (sigma 3-8 px,
amplitude 20-100, centres uniform over the frame); blob parameters are drawn
once per sequence.  With ``drift`` > 0 every blob moves by an independent
integer step in {-drift..drift}^2 per frame (a random walk); ``drift=0`` gives
a static sequence.  Noise: y = alpha * Poisson(x / alpha) + N(mu, sigma^2),
clipped at 0 (the reference API rejects negative data, pguresvt/svt.py:215)
and cast to float32.

Returned arrays are Fortran-ordered (rows, cols, frames), the layout the
reference binding hands to C++ without a copy (pguresvt/_pguresvt.pyx:72-93).
"""

import numpy as np


def blob_sequence(n, frames, blobs=8, drift=0, seed=0):
    """Clean phantom, float64, shape (n, n, frames), Fortran order."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cy = rng.uniform(0, n, blobs)
    cx = rng.uniform(0, n, blobs)
    sig = rng.uniform(3.0, 8.0, blobs)
    amp = rng.uniform(20.0, 100.0, blobs)
    if drift > 0:
        steps = rng.integers(-drift, drift + 1, size=(frames, blobs, 2))
        steps[0] = 0
        offs = np.cumsum(steps, axis=0)
    else:
        offs = np.zeros((frames, blobs, 2), dtype=np.int64)
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    out = np.empty((n, n, frames), dtype=np.float64, order="F")
    for f in range(frames):
        img = np.full((n, n), 10.0)
        for b in range(blobs):
            y0 = cy[b] + offs[f, b, 0]
            x0 = cx[b] + offs[f, b, 1]
            img += amp[b] * np.exp(-((yy - y0) ** 2 + (xx - x0) ** 2) / (2.0 * sig[b] ** 2))
        out[:, :, f] = img
    return out


def add_noise(clean, alpha=1.0, mu=0.0, sigma=8.0, seed=0):
    """Poisson-Gaussian noise, clipped at 0, float32, Fortran order."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    y = alpha * rng.poisson(clean / alpha) + rng.normal(mu, sigma, size=clean.shape)
    y = np.clip(y, 0.0, None)
    return np.asfortranarray(y.astype(np.float32))


def noisy_sequence(n, frames, blobs=8, drift=0, alpha=1.0, mu=0.0, sigma=8.0, seed=0):
    """(clean float64, noisy float32) pair, both Fortran-ordered."""
    clean = blob_sequence(n, frames, blobs=blobs, drift=drift, seed=seed)
    return clean, add_noise(clean, alpha=alpha, mu=mu, sigma=sigma, seed=seed)
