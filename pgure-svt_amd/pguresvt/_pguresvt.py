"""ctypes binding of libpgsvt.so: the pguresvt_{u8,u16,f,d} entry points.

Same names, keyword arguments, defaults and return value as the Cython module
of the reference (pguresvt/_pguresvt.pyx:151-384): each returns
(Y, estimates, result) where Y is the float64 output cube seen as a C-order
array of shape (n_frames, n_cols, n_rows) and estimates is float64 of shape
(4, n_frames) (rows: lambda, alpha, mu, sigma).

Deviation from the reference: pguresvt_f takes float32 input.  The reference
types it np.float_t (= float64) on the Python side but float on the C++ side,
so it can accept neither; here it is PGURESVT<float, double>.

The library runs the whole path on the GPU; there is no CPU fallback.  If the
library is missing or no gfx950 device is present, calls raise.
"""

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgsvt.so")

PGSVT_U8, PGSVT_U16, PGSVT_F32, PGSVT_F64 = 0, 1, 2, 3


class pgsvt_params(C.Structure):
    """Mirror of struct pgsvt_params (include/pgsvt.h)."""

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


_lib = None


def load():
    """Load libpgsvt.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libpgsvt.so not built: run `python pgure-svt_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        vp, u32, i64, dp = C.c_void_p, C.c_uint32, C.c_int64, C.POINTER(C.c_double)
        P = C.POINTER(pgsvt_params)
        L.pgsvt_last_error.restype = C.c_char_p
        L.pgsvt_version.restype = C.c_char_p
        L.pgsvt_default_params.argtypes = [P]
        L.pgsvt_denoise.argtypes = [vp, C.c_int, u32, u32, u32, P, vp, vp]
        L.pgsvt_window_trajectories.argtypes = [vp, C.c_int, u32, u32, u32, P, u32, vp]
        L.pgsvt_window_risk.argtypes = [vp, C.c_int, u32, u32, u32, P, u32, vp, vp, u32, vp]
        L.pgsvt_window_reconstruct.argtypes = [vp, C.c_int, u32, u32, u32, P, u32, C.c_double, vp]
        L.pgsvt_window_singular_values.argtypes = [vp, C.c_int, u32, u32, u32, P, u32, vp, C.POINTER(u32)]
        L.pgsvt_noise_estimate.argtypes = [vp, u32, u32, u32, u32, dp, dp, dp]
        L.pgsvt_delta2_signs.argtypes = [i64, C.c_uint64, vp]
        L.pgsvt_frame_leaves.argtypes = [vp, u32, C.c_int, u32, vp, vp, C.POINTER(u32)]
        _lib = L
    return _lib


class PGSVTError(RuntimeError):
    """Failure reported by libpgsvt (message from pgsvt_last_error)."""


def check(rc):
    if rc != 0:
        raise PGSVTError("pgsvt error %d: %s" % (rc, load().pgsvt_last_error().decode()))


def make_params(trajectory_length=15, patch_size=4, patch_overlap=1, motion_window=7, motion_filter=5,
                noise_method=4, max_iter=500, n_jobs=-1, random_seed=-1, optimize_pgure=True,
                exponential_weighting=True, motion_estimation=True, lambda1=0.0, noise_alpha=-1.0, noise_mu=-1.0,
                noise_sigma=-1.0, tol=1e-7):
    p = pgsvt_params()
    p.trajectory_length = trajectory_length
    p.patch_size = patch_size
    p.patch_overlap = patch_overlap
    p.motion_window = motion_window
    p.motion_filter = motion_filter
    p.noise_method = noise_method
    p.max_iter = max_iter
    p.n_jobs = n_jobs
    p.random_seed = random_seed
    p.optimize_pgure = int(bool(optimize_pgure))
    p.exponential_weighting = int(bool(exponential_weighting))
    p.motion_estimation = int(bool(motion_estimation))
    p.lambda1 = lambda1
    p.noise_alpha = noise_alpha
    p.noise_mu = noise_mu
    p.noise_sigma = noise_sigma
    p.tol = tol
    return p


def _run(X, dtype_code, np_dtype, kw):
    if not isinstance(X, np.ndarray) or X.ndim != 3 or X.dtype != np_dtype:
        raise ValueError("Buffer dtype mismatch, expected a 3-d %s array" % np.dtype(np_dtype).name)
    X = np.asfortranarray(X)
    nr, nc, nf = X.shape
    p = make_params(**kw)
    Y = np.empty((nr, nc, nf), dtype=np.float64, order="F")
    est = np.empty((nf, 4), dtype=np.float64, order="F")
    check(load().pgsvt_denoise(X.ctypes.data, dtype_code, nr, nc, nf, C.byref(p), Y.ctypes.data, est.ctypes.data))
    # the reference hands the column-major cube to numpy as a C-order array
    # of shape (slices, cols, rows) (pguresvt/_pguresvt.pyx:117-138), and the
    # column-major (n_frames x 4) matrix as shape (4, n_frames) (:98-109)
    Yv = Y.ravel(order="F").reshape((nf, nc, nr))
    Ev = est.ravel(order="F").reshape((4, nf))
    return Yv, Ev, 0


def _entry(code, np_dtype):
    def fn(input_images, trajectory_length=15, patch_size=4, patch_overlap=1, motion_window=7, motion_filter=5,
           noise_method=4, max_iter=500, n_jobs=-1, random_seed=-1, optimize_pgure=True,
           exponential_weighting=True, motion_estimation=True, lambda1=0.0, noise_alpha=-1.0, noise_mu=-1.0,
           noise_sigma=-1.0, tol=1e-7):
        kw = dict(trajectory_length=trajectory_length, patch_size=patch_size, patch_overlap=patch_overlap,
                  motion_window=motion_window, motion_filter=motion_filter, noise_method=noise_method,
                  max_iter=max_iter, n_jobs=n_jobs, random_seed=random_seed, optimize_pgure=optimize_pgure,
                  exponential_weighting=exponential_weighting, motion_estimation=motion_estimation,
                  lambda1=lambda1, noise_alpha=noise_alpha, noise_mu=noise_mu, noise_sigma=noise_sigma, tol=tol)
        return _run(input_images, code, np_dtype, kw)

    return fn


pguresvt_u8 = _entry(PGSVT_U8, np.uint8)
pguresvt_u16 = _entry(PGSVT_U16, np.uint16)
pguresvt_f = _entry(PGSVT_F32, np.float32)
pguresvt_d = _entry(PGSVT_F64, np.float64)

for _f, _n in ((pguresvt_u8, "pguresvt_u8"), (pguresvt_u16, "pguresvt_u16"), (pguresvt_f, "pguresvt_f"),
               (pguresvt_d, "pguresvt_d")):
    _f.__name__ = _n
    _f.__qualname__ = _n
