"""CPU oracle of the learned-lambda network (test infrastructure only).

FP64 forward and backward of the network n_in -> h -> h -> h -> 1 (GELU)
with the product operands rounded to bfloat16 (round to nearest even)
exactly as the GPU's matrix-core path rounds them; sums are exact-ish
FP64 here and FP32 on the GPU.  Parameter layout: W1 [n_in][h], b1, W2
[h][h], b2, W3 [h][h], b3, W4 [h][1], b4 (flat FP32).
"""

import math

import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gemm(A, B):
    return bf16(A) @ bf16(B)


_erf = np.vectorize(math.erf)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def unpack(params, n_in, h):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    shapes = [(n_in, h), (h,), (h, h), (h,), (h, h), (h,), (h, 1), (1,)]
    out = []
    for s in shapes:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    assert o == p.size
    return out


def forward(params, X, n_in, h):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params, n_in, h)
    p1 = gemm(X, W1) + b1
    a1 = gelu(p1)
    p2 = gemm(a1, W2) + b2
    a2 = gelu(p2)
    p3 = gemm(a2, W3) + b3
    a3 = gelu(p3)
    y = (gemm(a3, W4) + b4)[:, 0]
    return y, (X, p1, a1, p2, a2, p3, a3)


def gradient(params, X, t, n_in, h):
    """(flat gradient of mean((y - t)^2), loss) with the GPU's product rounding."""
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params, n_in, h)
    y, (X, p1, a1, p2, a2, p3, a3) = forward(params, X, n_in, h)
    n = X.shape[0]
    e = y - t
    loss = float(np.mean(e * e))
    dy = (2.0 * e / n)[:, None]
    gW4 = gemm(a3.T, dy)
    gb4 = dy.sum(0)
    da = gemm(dy, W4.T)
    grads = [gW4, gb4]
    for x, W, p in ((a2, W3, p3), (a1, W2, p2), (X, W1, p1)):
        dp = da * gelu_grad(p)
        grads = [gemm(x.T, dp), dp.sum(0)] + grads
        da = gemm(dp, W.T)
    return np.concatenate([g.ravel() for g in grads]), loss
