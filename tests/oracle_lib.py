"""ctypes access to the CPU FP64 oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests, smoke() and bench.py's CPU baseline use it.
"""

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

DTYPES = {np.dtype("uint8"): 0, np.dtype("uint16"): 1, np.dtype("float32"): 2, np.dtype("float64"): 3}


class Params(C.Structure):
    """Mirror of pgsvt_params (include/pgsvt.h)."""

    _fields_ = [
        ("trajectory_length", C.c_uint32),
        ("patch_size", C.c_uint32),
        ("patch_overlap", C.c_uint32),
        ("motion_window", C.c_uint32),
        ("motion_filter", C.c_int64),
        ("noise_method", C.c_uint32),
        ("max_iter", C.c_uint32),
        ("n_jobs", C.c_int64),
        ("random_seed", C.c_int64),
        ("optimize_pgure", C.c_int32),
        ("exponential_weighting", C.c_int32),
        ("motion_estimation", C.c_int32),
        ("lambda1", C.c_double),
        ("noise_alpha", C.c_double),
        ("noise_mu", C.c_double),
        ("noise_sigma", C.c_double),
        ("tol", C.c_double),
    ]


def make_params(**kw):
    d = dict(trajectory_length=15, patch_size=4, patch_overlap=1, motion_window=7, motion_filter=-1,
             noise_method=4, max_iter=500, n_jobs=-1, random_seed=101, optimize_pgure=1,
             exponential_weighting=1, motion_estimation=1, lambda1=-1.0, noise_alpha=-1.0,
             noise_mu=-1.0, noise_sigma=-1.0, tol=1e-7)
    d.update(kw)
    p = Params()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def geometry(p):
    nt = p.patch_size ** 2 - 1 if p.patch_size ** 2 < p.trajectory_length else p.trajectory_length
    fw = nt // 2
    return nt, fw, 2 * fw + 1


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(ORACLE_SO)
        vp, u32, i32, i64, u64, dp = C.c_void_p, C.c_uint32, C.c_int32, C.c_int64, C.c_uint64, C.POINTER(C.c_double)
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_svd_name.restype = C.c_char_p
        L.oracle_set_svd.argtypes = [C.c_int]
        L.oracle_denoise.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(Params), vp, vp]
        L.oracle_denoise_range.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(Params), u32, u32, vp, vp]
        L.oracle_window_trajectories.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(Params), u32, vp]
        L.oracle_window_risk.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(Params), u32, vp, vp, u32, vp]
        L.oracle_window_reconstruct.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(Params), u32, C.c_double, vp]
        L.oracle_window_singular_values.argtypes = [vp, C.c_int, u32, u32, u32, C.POINTER(Params), u32, vp,
                                                    C.POINTER(u32)]
        L.oracle_noise_estimate.argtypes = [vp, u32, u32, u32, u32, dp, dp, dp]
        L.oracle_pcg64.argtypes = [u64, u64, vp]
        L.oracle_median_filter.argtypes = [vp, C.c_int, u32, u32, u32, C.c_int, vp]
        L.oracle_hotpixel.argtypes = [vp, u32, u32, u32, C.c_double]
        L.oracle_perturbations.argtypes = [i64, u64, vp, vp]
        L.oracle_patch_indices.argtypes = [u32, u32, u32, vp]
        L.oracle_patch_indices.restype = u32
        L.oracle_noise_robust_mean.argtypes = [vp, u32, dp]
        L.oracle_noise_robust_var.argtypes = [vp, u32, dp]
        L.oracle_noise_mode.argtypes = [vp, u32, dp]
        L.oracle_noise_wls.argtypes = [vp, vp, u32, dp, dp]
        L.oracle_noise_split.argtypes = [vp, u32, C.POINTER(i32)]
        L.oracle_noise_tree.argtypes = [vp, u32, vp, u32]
        L.oracle_noise_tree.restype = u32
        FN = C.CFUNCTYPE(C.c_double, C.c_double, C.c_void_p)
        L.FN = FN
        L.oracle_sbplx.argtypes = [FN, vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_int, dp, dp, dp, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle error %d: %s" % (rc, lib().oracle_last_error().decode()))


def _xarg(X):
    X = np.asfortranarray(X)
    return X, DTYPES[X.dtype]


def set_svd(backend):
    """0 = FP64 Jacobi, 1 = LAPACK dgesdd; returns the backend name."""
    _check(lib().oracle_set_svd(backend))
    return lib().oracle_svd_name().decode()


def denoise(X, **kw):
    X, dt = _xarg(X)
    p = make_params(**kw)
    nr, nc, nf = X.shape
    Y = np.zeros((nr, nc, nf), order="F")
    est = np.zeros((nf, 4), order="F")
    _check(lib().oracle_denoise(X.ctypes.data, dt, nr, nc, nf, C.byref(p), Y.ctypes.data, est.ctypes.data))
    return Y, est


def denoise_range(X, t0, t1, **kw):
    X, dt = _xarg(X)
    p = make_params(**kw)
    nr, nc, nf = X.shape
    Y = np.zeros((nr, nc, nf), order="F")
    est = np.zeros((nf, 4), order="F")
    _check(lib().oracle_denoise_range(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t0, t1, Y.ctypes.data,
                                      est.ctypes.data))
    return Y, est


def window_trajectories(X, t, **kw):
    X, dt = _xarg(X)
    p = make_params(**kw)
    nr, nc, nf = X.shape
    nw = geometry(p)[2]
    pall = (nr - p.patch_size + 1) * (nc - p.patch_size + 1)
    out = np.zeros((nw, pall, 2), dtype=np.int32)
    _check(lib().oracle_window_trajectories(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, out.ctypes.data))
    return out


def window_risk(X, t, noise, lambdas, **kw):
    X, dt = _xarg(X)
    p = make_params(**kw)
    nr, nc, nf = X.shape
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    nz = np.ascontiguousarray(noise, dtype=np.float64)
    out = np.zeros(len(lam))
    _check(lib().oracle_window_risk(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, nz.ctypes.data, lam.ctypes.data,
                                    len(lam), out.ctypes.data))
    return out


def window_reconstruct(X, t, lam, **kw):
    X, dt = _xarg(X)
    p = make_params(**kw)
    nr, nc, nf = X.shape
    nw = geometry(p)[2]
    out = np.zeros((nr, nc, nw), order="F")
    _check(lib().oracle_window_reconstruct(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, lam, out.ctypes.data))
    return out


def window_singular_values(X, t, **kw):
    X, dt = _xarg(X)
    p = make_params(**kw)
    nr, nc, nf = X.shape
    nw = geometry(p)[2]
    pall = (nr - p.patch_size + 1) * (nc - p.patch_size + 1)
    out = np.zeros((pall, nw))
    n = C.c_uint32(0)
    _check(lib().oracle_window_singular_values(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, out.ctypes.data,
                                               C.byref(n)))
    return out[: n.value]


def noise_estimate(u, method=4, alpha=-1.0, mu=-1.0, sigma=-1.0):
    u = np.asfortranarray(u, dtype=np.float64)
    a, m, s = C.c_double(alpha), C.c_double(mu), C.c_double(sigma)
    nr, nc, nt = u.shape
    _check(lib().oracle_noise_estimate(u.ctypes.data, nr, nc, nt, method, C.byref(a), C.byref(m), C.byref(s)))
    return a.value, m.value, s.value


def pcg64(seed, n):
    out = np.zeros(n, dtype=np.uint64)
    _check(lib().oracle_pcg64(seed, n, out.ctypes.data))
    return out


def perturbations(seed, n):
    d1 = np.zeros(n, dtype=np.int8)
    d2 = np.zeros(n, dtype=np.float64)
    _check(lib().oracle_perturbations(seed, n, d1.ctypes.data, d2.ctypes.data))
    return d1, d2


def patch_indices(N, bs, step):
    n = lib().oracle_patch_indices(N, bs, step, None)
    out = np.zeros(n, dtype=np.uint32)
    lib().oracle_patch_indices(N, bs, step, out.ctypes.data)
    return out


def sbplx(f, start, lb, ub, step, ftol=1e-7, xtol_abs=1e-12, maxeval=500):
    L = lib()
    cb = L.FN(lambda x, _ctx: float(f(x)))
    xb, fb, xl = C.c_double(), C.c_double(), C.c_double()
    ne = C.c_int()
    rc = L.oracle_sbplx(cb, None, start, lb, ub, step, ftol, xtol_abs, maxeval, C.byref(xb), C.byref(fb),
                        C.byref(xl), C.byref(ne))
    return rc, xb.value, fb.value, xl.value, ne.value


def median_filter(X, r):
    X, dt = _xarg(X)
    nr, nc, nf = X.shape
    Z = np.zeros((nr, nc, nf), order="F")
    _check(lib().oracle_median_filter(X.ctypes.data, dt, nr, nc, nf, r, Z.ctypes.data))
    return Z


def hotpixel(seq, threshold):
    """HotPixelFilter of a uint16 cube (nr, nc, nf), returns a filtered copy."""
    out = np.asfortranarray(seq.astype(np.uint16, copy=True))
    nr, nc, nf = out.shape
    _check(lib().oracle_hotpixel(out.ctypes.data, nr, nc, nf, threshold))
    return out


# ---- the noise estimator's pieces (src/noise.hpp) ----
def _vec(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def noise_robust_mean(a):
    a = _vec(a)
    out = C.c_double()
    _check(lib().oracle_noise_robust_mean(a.ctypes.data, len(a), C.byref(out)))
    return out.value


def noise_robust_var(a):
    a = _vec(a)
    out = C.c_double()
    _check(lib().oracle_noise_robust_var(a.ctypes.data, len(a), C.byref(out)))
    return out.value


def noise_mode(a):
    a = _vec(a)
    out = C.c_double()
    _check(lib().oracle_noise_mode(a.ctypes.data, len(a), C.byref(out)))
    return out.value


def noise_wls(x, y):
    x, y = _vec(x), _vec(y)
    a, b = C.c_double(), C.c_double()
    _check(lib().oracle_noise_wls(x.ctypes.data, y.ctypes.data, len(x), C.byref(a), C.byref(b)))
    return a.value, b.value


def noise_split(block):
    b = np.asfortranarray(block, dtype=np.float64)
    out = C.c_int32()
    _check(lib().oracle_noise_split(b.ctypes.data, b.shape[0], C.byref(out)))
    return bool(out.value)


def noise_tree(frame):
    f = np.asfortranarray(frame, dtype=np.float64)
    n = lib().oracle_noise_tree(f.ctypes.data, f.shape[0], None, 0)
    out = np.zeros((n, 3), dtype=np.uint32)
    lib().oracle_noise_tree(f.ctypes.data, f.shape[0], out.ctypes.data, n)
    return out
