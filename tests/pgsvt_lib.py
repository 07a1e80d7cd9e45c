"""ctypes helpers around libpgsvt.so's window-level entry points for tests."""

import ctypes as C

import numpy as np

from pguresvt import _pguresvt as P
from oracle_lib import geometry

DT = {np.dtype("uint8"): 0, np.dtype("uint16"): 1, np.dtype("float32"): 2, np.dtype("float64"): 3}


def params(**kw):
    base = dict(trajectory_length=15, patch_size=4, patch_overlap=1, motion_window=7, motion_filter=-1,
                noise_method=4, max_iter=500, n_jobs=-1, random_seed=101, optimize_pgure=True,
                exponential_weighting=True, motion_estimation=True, lambda1=-1.0, noise_alpha=-1.0,
                noise_mu=-1.0, noise_sigma=-1.0, tol=1e-7)
    base.update(kw)
    return P.make_params(**base)


def _x(X):
    X = np.asfortranarray(X)
    return X, DT[X.dtype]


def denoise(X, **kw):
    X, dt = _x(X)
    nr, nc, nf = X.shape
    p = params(**kw)
    Y = np.zeros((nr, nc, nf), order="F")
    est = np.zeros((nf, 4), order="F")
    P.check(P.load().pgsvt_denoise(X.ctypes.data, dt, nr, nc, nf, C.byref(p), Y.ctypes.data, est.ctypes.data))
    return Y, est


def window_trajectories(X, t, **kw):
    X, dt = _x(X)
    nr, nc, nf = X.shape
    p = params(**kw)
    nw = geometry(p)[2]
    pall = (nr - p.patch_size + 1) * (nc - p.patch_size + 1)
    out = np.zeros((nw, pall, 2), dtype=np.int32)
    P.check(P.load().pgsvt_window_trajectories(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, out.ctypes.data))
    return out


def window_risk(X, t, noise, lambdas, **kw):
    X, dt = _x(X)
    nr, nc, nf = X.shape
    p = params(**kw)
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    nz = np.ascontiguousarray(noise, dtype=np.float64)
    out = np.zeros(len(lam))
    P.check(P.load().pgsvt_window_risk(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, nz.ctypes.data,
                                       lam.ctypes.data, len(lam), out.ctypes.data))
    return out


def window_reconstruct(X, t, lam, **kw):
    X, dt = _x(X)
    nr, nc, nf = X.shape
    p = params(**kw)
    nw = geometry(p)[2]
    out = np.zeros((nr, nc, nw), order="F")
    P.check(P.load().pgsvt_window_reconstruct(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, lam, out.ctypes.data))
    return out


def window_singular_values(X, t, **kw):
    X, dt = _x(X)
    nr, nc, nf = X.shape
    p = params(**kw)
    nw = geometry(p)[2]
    pall = (nr - p.patch_size + 1) * (nc - p.patch_size + 1)
    out = np.zeros((pall, nw))
    n = C.c_uint32(0)
    P.check(P.load().pgsvt_window_singular_values(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, out.ctypes.data,
                                                  C.byref(n)))
    return out[: n.value]


def noise_estimate(u, method=4, alpha=-1.0, mu=-1.0, sigma=-1.0):
    u = np.asfortranarray(u, dtype=np.float64)
    a, m, s = C.c_double(alpha), C.c_double(mu), C.c_double(sigma)
    nr, nc, nt = u.shape
    P.check(P.load().pgsvt_noise_estimate(u.ctypes.data, nr, nc, nt, method, C.byref(a), C.byref(m), C.byref(s)))
    return a.value, m.value, s.value


def frame_leaves(frame, on_gpu):
    """Robust (mean, variance) of the quadtree leaves of one normalised
    square frame, on the host or the GPU (pgsvt_frame_leaves)."""
    f = np.asfortranarray(frame, dtype=np.float64)
    n = f.shape[0]
    cnt = C.c_uint32(0)
    cap = (n // 8) ** 2 * 2 + 64
    mean, var = np.zeros(cap), np.zeros(cap)
    P.check(P.load().pgsvt_frame_leaves(f.ctypes.data, n, int(on_gpu), cap, mean.ctypes.data, var.ctypes.data,
                                        C.byref(cnt)))
    return mean[:cnt.value].copy(), var[:cnt.value].copy()


def delta2_signs(seed, n):
    out = np.zeros(n, dtype=np.uint8)
    P.check(P.load().pgsvt_delta2_signs(seed, n, out.ctypes.data))
    return out


def rel_l2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def set_svd_mapping(mode):
    """pgsvt_set_svd_mapping: 0 primary, 1 alternative, 2 autotuned (grid)."""
    P.check(P.load().pgsvt_set_svd_mapping(mode))
