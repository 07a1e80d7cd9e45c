"""Front end on the GPU: the hot-pixel kernel against the oracle (bit-exact
uint16) and the command-line program end to end."""

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from test_frontend import CLI, tiff_read, tiff_write

pytestmark = pytest.mark.gpu


def gpu_hotpixel(seq, threshold):
    from pguresvt import _pguresvt as P

    L = P.load()
    L.pgsvt_hotpixel_filter.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double]
    out = np.asfortranarray(seq.astype(np.uint16, copy=True))
    nr, nc, nf = out.shape
    P.check(L.pgsvt_hotpixel_filter(out.ctypes.data, nr, nc, nf, threshold))
    return out


def _hot_sequence(shape, seed, n_hot):
    rng = np.random.default_rng(seed)
    F = rng.normal(1000, 30, size=shape).clip(0, 65535).astype(np.uint16)
    nr, nc, nf = shape
    for k in range(nf):
        r = rng.integers(0, nr, n_hot)
        c = rng.integers(0, nc, n_hot)
        F[r, c, k] = rng.integers(20000, 65535, n_hot)
        F[nr // 2:nr // 2 + 3, nc // 2:nc // 2 + 3, k] = 50000  # a block: order-dependent replacement
    return F


@pytest.mark.parametrize("shape,thr", [((256, 256, 3), 6.0), ((33, 33, 2), 4.0), ((64, 48, 2), 0.0),
                                       ((17, 40, 1), 10.0)])
def test_hotpixel_matches_oracle(gpu, shape, thr):
    F = _hot_sequence(shape, sum(shape), 40)
    assert np.array_equal(gpu_hotpixel(F, thr), O.hotpixel(F, thr))


def test_hotpixel_known_answers(gpu):
    F = np.zeros((4, 2, 1), dtype=np.uint16)
    F[:, 0, 0] = 10
    F[:, 1, 0] = 13
    F[0, 0, 0] = 1000
    assert gpu_hotpixel(F, 3.0)[0, 0, 0] == 12
    b, H = 100, 1000
    G = np.full((12, 12, 1), b, dtype=np.uint16)
    G[5:8, 5:8, 0] = H
    out = gpu_hotpixel(G, 0.0)
    assert np.array_equal(out[5:8, 5:8, 0], [[b, 325, b], [550, 775, 212], [b, 325, b]])


def test_cli_end_to_end(gpu, tmp_path):
    # pgure-svt on a 16-bit TIFF with the hot-pixel filter and a fixed lambda:
    # the written frames equal hot-pixel filter -> pgsvt_denoise (uint16 in)
    # -> truncation to uint16, done through the library
    from pguresvt import _pguresvt as P

    X = np.load(os.path.join(os.path.dirname(__file__), "golden", "data.npz"))["a"]  # (32, 32, 16) uint16
    rng = np.random.RandomState(101)
    Y = (X + 100.0 + 100.0 * rng.randn(*X.shape)).clip(0, None).astype(np.uint16)
    Y[3, 4, 2] = 60000
    tiff_write(str(tmp_path / "seq.tif"), Y)
    params = ("filename : seq.tif\nstart_frame : 2\nend_frame : 16\nhot_pixel : 5\n"
              "optimize_pgure : false\nlambda : 2.0\nmedian_filter : 0\nrandom_seed : 9\n"
              "trajectory_length : 15\npatch_size : 4\n")
    (tmp_path / "p.txt").write_text(params)
    r = subprocess.run([CLI, str(tmp_path / "p.txt")], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "seq-CLEANED.tif" in r.stderr
    out, bits = tiff_read(str(tmp_path / "seq-CLEANED.tif"))
    assert bits == 16 and out.shape == (32, 32, 15)

    seq = gpu_hotpixel(Y[:, :, 1:16], 5.0)
    L = P.load()
    p = P.make_params(trajectory_length=15, patch_size=4, patch_overlap=1, motion_window=7, motion_filter=0,
                      noise_method=4, max_iter=1000, n_jobs=-1, random_seed=9, optimize_pgure=0,
                      exponential_weighting=1, motion_estimation=1, lambda1=2.0, noise_alpha=-1.0,
                      noise_mu=-1.0, noise_sigma=-1.0, tol=1e-7)
    Yd = np.zeros(seq.shape, order="F")
    est = np.zeros((15, 4), order="F")
    P.check(L.pgsvt_denoise(seq.ctypes.data, P.PGSVT_U16, 32, 32, 15, C.byref(p), Yd.ctypes.data, est.ctypes.data))
    expect = np.where(Yd > 0, np.minimum(Yd, 65535), 0).astype(np.uint16)
    assert np.array_equal(out, expect)
