"""pgsvt_denoise_device_noalloc (the entry point that allocates no device
memory) against pgsvt_denoise_device on the same inputs: two calls in a
row, the second on a smaller sequence, with the PGURE search and with a
fixed lambda.  The kernels and the order of every sum are shared, so the
frames and estimates must be identical."""

import ctypes as C
import sys

import numpy as np
import pytest

import pgsvt_lib as G
from pguresvt import _pguresvt as P
from pguresvt import phantom

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class _Hip:
    def __init__(self):
        self.L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), n) == 0
        return p.value

    def h2d(self, d, a):
        assert self.L.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, d):
        assert self.L.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0


def _run(entry, X, **kw):
    L = P.load()
    # the entry point runs in FP32 with the default SVD mapping, whatever
    # earlier tests of the process selected
    L.pgsvt_set_precision.argtypes = [C.c_int]
    P.check(L.pgsvt_set_precision(32))
    P.check(L.pgsvt_set_svd_mapping(0))
    hip = _Hip()
    X = np.asfortranarray(X.astype(np.float32))
    nr, nc, nf = X.shape
    p = G.params(**kw)
    xd, yd = hip.malloc(X.nbytes), hip.malloc(8 * X.size)
    try:
        hip.h2d(xd, X)
        est = np.zeros((nf, 4), order="F")
        Y = np.zeros(X.shape, order="F")
        if entry == "noalloc":
            f = L.pgsvt_denoise_device_noalloc
            f.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P.pgsvt_params),
                          C.c_void_p, C.c_void_p, C.c_void_p]
            P.check(f(xd, 2, nr, nc, nf, C.byref(p), yd, est.ctypes.data, None))
        else:
            f = L.pgsvt_denoise_device
            f.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P.pgsvt_params),
                          C.c_void_p, C.c_void_p]
            P.check(f(xd, 2, nr, nc, nf, C.byref(p), yd, est.ctypes.data))
        hip.d2h(Y, yd)
    finally:
        hip.L.hipFree(xd)
        hip.L.hipFree(yd)
    return Y, est


@pytest.mark.parametrize("kw", [dict(), dict(optimize_pgure=False, lambda1=2.0), dict(motion_filter=3)])
def test_noalloc_matches_device_entry(gpu, kw):
    big = phantom.noisy_sequence(64, 20, drift=2, alpha=1.0, sigma=8.0, seed=2)[1]
    small = phantom.noisy_sequence(48, 16, drift=1, alpha=0.5, sigma=5.0, seed=9)[1]
    for name, X in (("64x64x20", big), ("48x48x16", small)):  # second call smaller
        log("noalloc vs device", name, kw)
        Yn, en = _run("noalloc", X, **kw)
        Yd, ed = _run("device", X, **kw)
        assert np.array_equal(en, ed), (en[:3], ed[:3])
        assert np.array_equal(Yn, Yd), float(np.abs(Yn - Yd).max())
        assert np.isfinite(Yn).all() and np.abs(Yn).max() > 0


def test_noalloc_rejects_oversized_shape(gpu):
    # a 256 x 256 window with PGURE needs far more than the 64 MiB array
    L = P.load()
    X = np.zeros((256, 256, 15), np.float32, order="F")
    f = L.pgsvt_denoise_device_noalloc
    f.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(P.pgsvt_params),
                  C.c_void_p, C.c_void_p, C.c_void_p]
    p = G.params()
    est = np.zeros((15, 4), order="F")
    rc = f(X.ctypes.data, 2, 256, 256, 15, C.byref(p), X.ctypes.data, est.ctypes.data, None)
    assert rc == -1 and b"bytes of device scratch" in L.pgsvt_last_error()
