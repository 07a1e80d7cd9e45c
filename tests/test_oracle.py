"""The CPU FP64 oracle checked against what pins it: numpy's PCG64, closed
forms, LAPACK, brute force, and the reference test suite's own thresholds."""

import numpy as np
import pytest

import oracle_lib as O
from pguresvt import phantom

INC = (6364136223846793005 << 64) | 1442695040888963407
MULT = (2549297995355413924 << 64) | 4865540595714422341


def numpy_pcg64(seed):
    g = np.random.PCG64()
    st = ((seed + INC) * MULT + INC) % (1 << 128)
    g.state = {"bit_generator": "PCG64", "state": {"state": st, "inc": INC}, "has_uint32": 0, "uinteger": 0}
    return g


@pytest.mark.parametrize("seed", [0, 1, 101, 2**40 + 7])
def test_pcg64_matches_numpy(seed):
    # pcg64 seeding and XSL-RR output (src/pcg/pcg_random.hpp) reproduce numpy's PCG64
    assert np.array_equal(numpy_pcg64(seed).random_raw(4000).astype(np.uint64), O.pcg64(seed, 4000))


def test_perturbations_closed_form():
    # delta1 from the first N draws (bernoulli 0.5 -> -1), delta2 from the next N
    # (bernoulli vP -> -sqrt(vQ/vP), else sqrt(vP/vQ)); libstdc++'s bernoulli is
    # u = double(draw) * 2^-64 < p (src/pgure.hpp:167-186)
    n = 5000
    d1, d2 = O.perturbations(101, n)
    raw = numpy_pcg64(101).random_raw(2 * n).astype(np.uint64)
    u = raw.astype(np.float64) * 2.0**-64
    vP = 0.5 + 0.5 / np.sqrt(5.0)
    vQ = 1 - vP
    assert np.array_equal(d1, np.where(u[:n] < 0.5, -1, 1))
    assert np.array_equal(d2, np.where(u[n:] < vP, -np.sqrt(vQ / vP), np.sqrt(vP / vQ)))
    # E[delta2] = 0 and E[delta2^2] = 1 by construction
    assert abs(vP * -np.sqrt(vQ / vP) + vQ * np.sqrt(vP / vQ)) < 1e-15


@pytest.mark.parametrize("N,bs,step,count", [(64, 4, 1, 3721), (256, 4, 1, 64009), (32, 4, 1, 841),
                                             (1024, 8, 2, 260096)])
def test_patch_index_counts(N, bs, step, count):
    # index set i*nyMbs + j of src/svt.hpp:61-92, including the edges
    idx = O.patch_indices(N, bs, step)
    assert len(idx) == count
    n = N - bs
    brute = set()
    for i in range(0, n + 1, step):
        for j in range(0, n + 1, step):
            brute.add(i * n + j)
    for i in range(0, n + 1, step):
        brute.add((n + 1) * i + n)
        brute.add((n + 1) * n + i)
    assert np.array_equal(idx, np.array(sorted(brute), dtype=np.uint32))


def test_svd_backends_agree():
    clean, noisy = phantom.noisy_sequence(32, 15, drift=1, seed=5)
    O.set_svd(0)
    s0 = O.window_singular_values(noisy, 7, motion_filter=-1)
    try:
        O.set_svd(1)
    except RuntimeError:
        pytest.skip("no LAPACK")
    s1 = O.window_singular_values(noisy, 7, motion_filter=-1)
    O.set_svd(0)
    assert np.max(np.abs(s0 - s1)) < 1e-12 * np.max(s0)
    # and numpy's LAPACK on the reference-frame Casorati matrix of patch 0
    u = noisy[:, :, 0:15].astype(np.float64) / noisy[:, :, 0:15].max()
    tr = O.window_trajectories(noisy, 7, motion_filter=-1)
    cas = np.stack([u[tr[k, 0, 0]:tr[k, 0, 0] + 4, tr[k, 0, 1]:tr[k, 0, 1] + 4, k].ravel(order="F")
                    for k in range(15)], axis=1)
    assert np.allclose(np.linalg.svd(cas, compute_uv=False), s0[0], rtol=1e-12, atol=1e-14)


def test_reconstruct_identity():
    # plain SVT at lambda = 0 keeps every singular value: the overlap-add of the
    # exact patches returns the input wherever a patch covers it (src/svt.hpp:121-167)
    clean, noisy = phantom.noisy_sequence(32, 15, drift=1, seed=3)
    vol = O.window_reconstruct(noisy, 7, 0.0, motion_filter=-1, exponential_weighting=0)
    u = noisy[:, :, 0:15].astype(np.float64) / noisy[:, :, 0:15].max()
    ref = vol[:, :, 7]
    assert np.allclose(ref, u[:, :, 7], rtol=0, atol=1e-12)


def test_reconstruct_motion_off_zero_pattern():
    # motion off: only the reference frame has real positions, every other frame
    # sees all patches at (0,0), so only its top-left block is non-zero
    clean, noisy = phantom.noisy_sequence(32, 15, drift=0, seed=4)
    vol = O.window_reconstruct(noisy, 7, 1.0, motion_filter=-1, motion_estimation=0)
    for k in range(15):
        if k == 7:
            continue
        m = np.zeros((32, 32), bool)
        m[:4, :4] = True
        assert np.all(vol[:, :, k][~m] == 0)
        assert np.any(vol[:, :, k][m] != 0)


def coverage(tr, shape, bs, sel=None):
    """Brute-force coverage mask of the patch positions tr[k, p] = (row, col)."""
    nr, nc, nw = shape
    cov = np.zeros(shape, bool)
    idx = range(tr.shape[1]) if sel is None else sel
    for k in range(nw):
        for p in idx:
            r, c = tr[k, p]
            cov[r:r + bs, c:c + bs, k] = True
    return cov


def test_risk_identity_closed_form():
    # plain SVT at lambda = 0 returns every input exactly where a patch covers
    # it and 0 elsewhere, so the second difference vanishes and, with the
    # delta1 term identically zero,
    #   R = (1/N)[sum_uncov U^2 - (alpha+mu_p) sum U + 2 mu_p sum_cov U + mu_p] - sigma_p^2
    # with mu_p = sigma, sigma_p = mu (constructor arguments swapped)
    clean, noisy = phantom.noisy_sequence(32, 15, drift=1, seed=8)
    alpha, mu, sigma = 0.02, 0.03, 0.05
    r = O.window_risk(noisy, 7, [alpha, mu, sigma], [0.0], motion_filter=-1, exponential_weighting=0)[0]
    u = noisy[:, :, 0:15].astype(np.float64) / noisy[:, :, 0:15].max()
    tr = O.window_trajectories(noisy, 7, motion_filter=-1)
    cov = coverage(tr, u.shape, 4)
    assert not cov.all()  # the drift leaves some voxels uncovered
    N = u.size
    mup, sigp = sigma, mu
    expect = ((u[~cov] ** 2).sum() - (alpha + mup) * u.sum() + 2 * mup * u[cov].sum() + mup) / N - sigp**2
    assert abs(r - expect) < 1e-10 * abs(expect)
    # a wrong sign on the mu_p sums or an unswapped (mu, sigma) moves it visibly
    wrong = ((u[~cov] ** 2).sum() - (alpha + mu) * u.sum() + 2 * mu * u[cov].sum() + mu) / N - sigma**2
    assert abs(r - wrong) > 1e-6


def sbplx_quad(c, start, ub=100.0):
    calls = []

    def f(x):
        calls.append(x)
        return (x - c) ** 2 + 1.0

    rc, xb, fb, xl, ne = O.sbplx(f, start, 0.0, ub, np.sqrt(start))
    return rc, xb, fb, xl, ne, calls


def test_sbplx_minimises_and_returns_last():
    rc, xb, fb, xl, ne, calls = sbplx_quad(3.3, 0.7)
    assert rc in (3, 4)
    assert abs(xb - 3.3) < 1e-3
    assert xl == calls[-1] and ne == len(calls)
    # bounds respected
    rc, xb, fb, xl, ne, calls = sbplx_quad(-5.0, 0.4)
    assert min(calls) >= 0.0 and xb < 1e-6
    rc, xb, fb, xl, ne, calls = sbplx_quad(500.0, 2.0)
    assert max(calls) <= 100.0 and abs(xb - 100.0) < 1e-6


def test_median_filter_bruteforce():
    rng = np.random.default_rng(3)
    X = np.asfortranarray(rng.uniform(0, 300, size=(20, 20, 2)).astype(np.float32))
    Z = O.median_filter(X, 2)
    F = np.trunc(X.astype(np.float64)).astype(np.int64)
    for f in range(2):
        P = np.pad(F[:, :, f], 2, mode="edge")
        for r in range(20):
            for c in range(20):
                assert Z[r, c, f] == np.median(P[r:r + 5, c:c + 5])


def nsed(A, B):
    A_m = A - A.mean()
    B_m = B - B.mean()
    return 0.5 * np.linalg.norm(A_m - B_m) ** 2 / (np.linalg.norm(A_m) ** 2 + np.linalg.norm(B_m) ** 2)


def test_reference_quality_threshold_fixture():
    # pguresvt/tests/test_svt.py:61-90 on the reference's own fixture data.npz
    import os
    X = np.load(os.path.join(os.path.dirname(__file__), "golden", "data.npz"))["a"]
    rng = np.random.RandomState(101)
    Y = X + 100.0 + 100.0 * rng.randn(*X.shape)
    Y[Y < 0.0] = 0.0
    Y = Y.astype(np.uint16)
    Z, est = O.denoise(Y, random_seed=101, motion_filter=5, n_jobs=4)
    assert nsed(X, Z) < 0.025


def _smooth_texture(n, seed):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    return gaussian_filter(rng.standard_normal((n, n)), 2.0, mode="wrap") * 50 + 100


@pytest.mark.parametrize("dy,dx", [(1, 2), (-2, 1), (0, -3), (3, 0)])
@pytest.mark.parametrize("t", [0, 7, 15])
def test_arps_recovers_known_shift(dy, dx, t):
    # A smooth texture translated by (dy, dx) per frame.  ARPS moves every
    # block from its grid position in the source frame (src/arps.hpp:186-190,
    # 370-373), so frame k of the trajectory is the grid position plus one
    # shift forward of the reference frame and minus one shift behind it.
    # ARPS is a local search, so (1) most interior blocks land exactly there
    # and (2) every block ends on a rood local minimum of its cost inside the
    # search window: costs only fall along the LDSP/SDSP path and ties keep
    # the centre (src/arps.hpp:289-294, 299-368).
    n, nf, bs, mw = 48, 17, 4, 7
    base = _smooth_texture(n, 100 + dy * 7 + dx)
    X = np.asfortranarray(np.stack([np.roll(base, (dy * k, dx * k), axis=(0, 1)) for k in range(nf)], axis=2))
    tr = O.window_trajectories(X, t, motion_estimation=1, motion_window=mw)
    n1 = n - bs + 1
    s0, ref = (0, t) if t < 7 else ((t - 7, 7) if t < nf - 7 else (nf - 15, t - (nf - 15)))
    u = X[:, :, s0:s0 + 15] / X[:, :, s0:s0 + 15].max()
    it = np.arange(n1 * n1)
    gi, gj = it % n1, it // n1
    inner = (gi >= 8) & (gi <= n1 - 9) & (gj >= 8) & (gj <= n1 - 9)
    for k in range(15):
        if k == ref:
            assert (tr[k, :, 0] == gi).all() and (tr[k, :, 1] == gj).all()
            continue
        sgn = 1 if k > ref else -1
        ok = (tr[k, inner, 0] == gi[inner] + sgn * dy) & (tr[k, inner, 1] == gj[inner] + sgn * dx)
        assert ok.mean() > 0.75, (k, ref, ok.mean())
        src = k - 1 if k > ref else k + 1  # the block comes from the neighbouring frame towards ref
        for b in np.where(inner)[0][::7]:
            i, j = gi[b], gj[b]
            A = u[i:i + bs, j:j + bs, src]
            cost = lambda y, x: ((A - u[y:y + bs, x:x + bs, k]) ** 2).sum() / (bs * bs)
            y, x = tr[k, b]
            c0 = cost(y, x)
            for ny, nx in ((y - 1, x), (y + 1, x), (y, x - 1), (y, x + 1)):
                if 0 <= ny < n1 and 0 <= nx < n1 and abs(ny - i) <= mw and abs(nx - j) <= mw:
                    assert c0 <= cost(ny, nx) + 1e-12, (k, b, (y, x), (ny, nx))


def test_arps_motion_window_clamps_search():
    # a shift beyond the search window cannot be reached: no interior block
    # may report a displacement larger than motion_window (src/arps.hpp:299-368)
    n, nf, bs, mw = 48, 16, 4, 3
    base = _smooth_texture(n, 11)
    X = np.asfortranarray(np.stack([np.roll(base, (5 * k, 0), axis=(0, 1)) for k in range(nf)], axis=2))
    tr = O.window_trajectories(X, 7, motion_estimation=1, motion_window=mw)
    n1 = n - bs + 1
    it = np.arange(n1 * n1)
    gi, gj = it % n1, it // n1
    for k in range(15):
        assert np.abs(tr[k, :, 0] - gi).max() <= mw and np.abs(tr[k, :, 1] - gj).max() <= mw
