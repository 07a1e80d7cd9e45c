"""PGURE over a lambda grid, one captured graph per batch of windows, and
batches of independent sequences of uneven shapes (no reference
counterpart: the reference searches lambda with SBPLX per window,
src/pgure.hpp:196-237)."""

import ctypes as C

import numpy as np

from . import _pguresvt as _P


class Job(C.Structure):
    """Mirror of pgsvt_job (include/pgsvt.h)."""

    _fields_ = [("X", C.c_void_p), ("dtype", C.c_int), ("n_rows", C.c_uint32), ("n_cols", C.c_uint32),
                ("n_frames", C.c_uint32), ("params", _P.pgsvt_params), ("Y", C.c_void_p),
                ("estimates", C.c_void_p), ("status", C.c_int)]


def _lib():
    L = _P.load()
    if not getattr(L, "_batch_bound", False):
        vp, u32 = C.c_void_p, C.c_uint32
        L.pgsvt_denoise_grid.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(_P.pgsvt_params), u32, vp, vp]
        L.pgsvt_denoise_batch.argtypes = [C.POINTER(Job), u32, u32, u32]
        L._batch_bound = True
    return L


_DT = {np.dtype("uint8"): 0, np.dtype("uint16"): 1, np.dtype("float32"): 2, np.dtype("float64"): 3}


def denoise_grid(X, n_lambda=64, **params):
    """Denoise X (rows, cols, frames) choosing each window's lambda as the
    least PGURE risk over n_lambda grid candidates: (Y, estimates (frames, 4))."""
    X = np.asfortranarray(X)
    nr, nc, nf = X.shape
    p = _P.make_params(**params)
    Y = np.zeros((nr, nc, nf), order="F")
    est = np.zeros((nf, 4), order="F")
    _P.check(_lib().pgsvt_denoise_grid(X.ctypes.data, _DT[X.dtype], nr, nc, nf, C.byref(p), n_lambda, Y.ctypes.data,
                                       est.ctypes.data))
    return Y, est


def denoise_batch(items, n_lambda=64, concurrency=4):
    """items: list of (X, params dict).  Returns [(Y, estimates)] in order."""
    xs, outs, jobs = [], [], (Job * len(items))()
    for k, (X, params) in enumerate(items):
        X = np.asfortranarray(X)
        nr, nc, nf = X.shape
        Y = np.zeros((nr, nc, nf), order="F")
        est = np.zeros((nf, 4), order="F")
        xs.append(X)
        outs.append((Y, est))
        jobs[k] = Job(X.ctypes.data, _DT[X.dtype], nr, nc, nf, _P.make_params(**params), Y.ctypes.data,
                      est.ctypes.data, 0)
    _P.check(_lib().pgsvt_denoise_batch(jobs, len(items), n_lambda, concurrency))
    return outs
