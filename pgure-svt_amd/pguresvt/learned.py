"""Learned lambda proposal (no reference counterpart).

The reference searches lambda for every window with SBPLX over the PGURE
risk (src/pgure.hpp:196-237).  Here a small network proposes it instead:
log-scaled sorted singular values of 256 fixed-sample patches of the window
plus the window's estimated noise parameters (alpha, mu, sigma) go through
n_in -> 128 -> 128 -> 128 -> 1 (GELU) to log(lambda).  The network runs on
the GPU's BF16 matrix cores with FP32 accumulation, and is trained there
with the library's own forward/backward kernels on (features, PGURE-optimal
lambda) pairs produced by the GPU PGURE search itself.
"""

import ctypes as C
import math

import numpy as np

from . import _pguresvt as _P

HIDDEN = 128
N_SAMPLE = 256  # sampled patches per window


def _lib():
    L = _P.load()
    if not getattr(L, "_learned_bound", False):
        fp, u32, u64 = C.POINTER(C.c_float), C.c_uint32, C.c_uint64
        L.pgsvt_mlp_forward.argtypes = [C.c_void_p, u32, u32, C.c_void_p, u32, C.c_void_p]
        L.pgsvt_mlp_gradient.argtypes = [C.c_void_p, u32, u32, C.c_void_p, C.c_void_p, u32, C.c_void_p, fp]
        L.pgsvt_mlp_train.argtypes = [C.c_void_p, C.c_void_p, u32, u32, u32, u32, u32, C.c_float, u64, C.c_void_p,
                                      C.c_void_p]
        L._learned_bound = True
    return L


def param_count(n_in, h=HIDDEN):
    return n_in * h + h + 2 * (h * h + h) + h + 1


def init_params(n_in, h=HIDDEN, seed=0):
    """He-scaled normal weights, zero biases (float32, the library's layout:
    W1 [n_in][h], b1, W2 [h][h], b2, W3 [h][h], b3, W4 [h][1], b4)."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((n_in, h), (h, h), (h, h), (h, 1)):
        parts.append(rng.standard_normal(fan_in * fan_out) * math.sqrt(2.0 / fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


def mlp_forward(params, X, n_in, h=HIDDEN):
    X = np.ascontiguousarray(X, dtype=np.float32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    out = np.zeros(X.shape[0], dtype=np.float32)
    _P.check(_lib().pgsvt_mlp_forward(params.ctypes.data, n_in, h, X.ctypes.data, X.shape[0], out.ctypes.data))
    return out


def mlp_gradient(params, X, t, n_in, h=HIDDEN):
    X = np.ascontiguousarray(X, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    params = np.ascontiguousarray(params, dtype=np.float32)
    g = np.zeros_like(params)
    loss = C.c_float()
    _P.check(_lib().pgsvt_mlp_gradient(params.ctypes.data, n_in, h, X.ctypes.data, t.ctypes.data, X.shape[0],
                                       g.ctypes.data, C.byref(loss)))
    return g, loss.value


def mlp_train(X, t, params, epochs, batch=256, lr=1e-3, seed=0, h=HIDDEN):
    """Adam on the mean squared error; returns (params, loss per epoch)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    losses = np.zeros(epochs, dtype=np.float32)
    _P.check(_lib().pgsvt_mlp_train(X.ctypes.data, t.ctypes.data, X.shape[0], X.shape[1], h, epochs, batch, lr, seed,
                                    p.ctypes.data, losses.ctypes.data))
    return p, losses
