"""Independent pins of the CPU FP64 oracle's risk, thresholding, ARPS and
noise-estimator pieces: numpy brute force written from the reference text,
closed forms and hand-traced inputs.  Each check names the mistake it
catches."""

import numpy as np
import pytest

import oracle_lib as O

INC = (6364136223846793005 << 64) | 1442695040888963407
MULT = (2549297995355413924 << 64) | 4865540595714422341


def numpy_pcg64(seed):
    g = np.random.PCG64()
    st = ((seed + INC) * MULT + INC) % (1 << 128)
    g.state = {"bit_generator": "PCG64", "state": {"state": st, "inc": INC}, "has_uint32": 0, "uinteger": 0}
    return g


# ----------------------------------------------------------------------------
# Brute-force PGURE risk (src/pgure.hpp:26-137, src/svt.hpp:58-167)
# ----------------------------------------------------------------------------

def bf_patch_set(N, bs, step):
    """Patch index set of src/svt.hpp:61-92 as a Python set union."""
    n = N - bs
    s = {i * n + j for i in range(0, n + 1, step) for j in range(0, n + 1, step)}
    s |= {(n + 1) * i + n for i in range(0, n + 1, step)}
    s |= {(n + 1) * n + i for i in range(0, n + 1, step)}
    return sorted(s)


def bf_reconstruct(u, traj, idx, bs, lam, exp_w):
    """SVT<T>::Reconstruct with numpy's LAPACK SVD per patch
    (src/svt.hpp:99-116, 121-167): B = U diag(s') V^T, overlap-added with
    unit weights, divided by the weights, non-finite -> 0."""
    nr, nc, nt = u.shape
    v = np.zeros_like(u)
    w = np.zeros_like(u)
    for it in idx:
        A = np.stack([u[traj[k, it, 0]:traj[k, it, 0] + bs, traj[k, it, 1]:traj[k, it, 1] + bs, k].ravel(order="F")
                      for k in range(nt)], axis=1)
        U, s, Vt = np.linalg.svd(A, full_matrices=False)
        thr = np.abs(s.max() * np.exp(-0.5 * lam * s * s)) if exp_w else lam
        st = np.sign(s) * np.maximum(np.abs(s) - thr, 0.0)
        B = U @ np.diag(st) @ Vt
        for k in range(nt):
            r, c = traj[k, it]
            v[r:r + bs, c:c + bs, k] += B[:, k].reshape(bs, bs, order="F")
            w[r:r + bs, c:c + bs, k] += 1.0
    with np.errstate(invalid="ignore", divide="ignore"):
        out = v / w
    out[~np.isfinite(out)] = 0.0
    return out


def bf_risk_terms(u, traj, idx, bs, lam, exp_w, seed, alpha, mu_drv, sigma_drv):
    """The terms of CalculatePGURE (src/pgure.hpp:120-137) for the PGURE
    object the driver builds: the constructor receives (alpha, sigma, mu)
    (src/pguresvt.hpp:133), so inside the risk mu = the driver's sigma and
    sigma = the driver's mu.  delta1 is an integer cube, so delta1 * eps1
    multiplies by int(eps1) = 0 (src/pgure.hpp:80, 156)."""
    mu, sigma = sigma_drv, mu_drv
    sigmasq = sigma * sigma
    N = u.size
    eps1 = u.max() * 0.0001
    eps2 = 100 * eps1
    raw = numpy_pcg64(seed).random_raw(2 * N).astype(np.uint64)
    c = raw.astype(np.float64) * 2.0**-64
    vP = 0.5 + 0.5 * 1.0 / np.sqrt(1.0 * 1.0 + 4)
    vQ = 1 - vP
    d1 = np.where(c[:N] < 0.5, -1, 1).reshape(u.shape, order="F")
    d2 = np.where(c[N:] < vP, -1 * np.sqrt(vQ / vP), np.sqrt(vP / vQ)).reshape(u.shape, order="F")
    U1in = u + d1 * int(eps1)
    Uh = bf_reconstruct(u, traj, idx, bs, lam, exp_w)
    U1 = bf_reconstruct(U1in, traj, idx, bs, lam, exp_w)
    U2p = bf_reconstruct(u + d2 * eps2, traj, idx, bs, lam, exp_w)
    U2m = bf_reconstruct(u - d2 * eps2, traj, idx, bs, lam, exp_w)
    t = {
        "fit": (np.abs(Uh - u) ** 2).sum(),
        "offset": -(alpha + mu) * u.sum(),
        "div1": 2 / eps1 * (d1 * (alpha * u - alpha * mu + sigmasq) * (U1 - Uh)).sum(),
        "div2": -(2 * sigmasq * alpha / (eps2 * eps2) * (d2 * (U2p - 2 * Uh + U2m)).sum()),
        "mean": 2 * mu * Uh.sum(),
        "mu": mu,
    }
    risk = (1.0 / N) * (t["fit"] + t["offset"] + t["div1"] + t["div2"] + t["mean"] + t["mu"]) - sigmasq
    return risk, t, N, sigmasq


def motion_off_traj(nr, nw, ref, bs):
    """Q2: with motion estimation off only the reference frame holds the grid
    positions (row = it % n1, col = it / n1); every other frame keeps (0, 0)
    (src/arps.hpp:42, 60-64)."""
    n1 = nr - bs + 1
    tr = np.zeros((nw, n1 * n1, 2), dtype=np.int64)
    it = np.arange(n1 * n1)
    tr[ref, :, 0] = it % n1
    tr[ref, :, 1] = it // n1
    return tr


def tiny_sequence(n, nf, seed):
    rng = np.random.default_rng(seed)
    base = rng.uniform(10, 60, size=(n, n))
    X = np.stack([np.roll(base, (k, -k), axis=(0, 1)) + rng.normal(0, 6, size=(n, n)) for k in range(nf)], axis=2)
    return np.asfortranarray(np.clip(X, 0, None))


@pytest.mark.parametrize("motion", [0, 1])
@pytest.mark.parametrize("exp_w", [1, 0])
@pytest.mark.parametrize("step", [1, 2])
def test_risk_bruteforce_numpy(motion, exp_w, step):
    # 8 x 8 x 3 frames, patch 2 (Nt = 3: bs^2 = 4 is not below 3), window
    # t = 1 (frames 0-2, reference 1).  Motion off: trajectories from Q2;
    # motion on: the oracle's ARPS trajectories (pinned by the ARPS tests
    # below).  Every term is formed literally, with numpy's LAPACK SVD.
    n, nf, bs, t, seed = 8, 3, 2, 1, 77
    X = tiny_sequence(n, nf, 5 + step)
    kw = dict(trajectory_length=3, patch_size=bs, patch_overlap=step, motion_filter=-1, motion_estimation=motion,
              exponential_weighting=exp_w, random_seed=seed, motion_window=3)
    u = X / X.max()
    traj = motion_off_traj(n, 3, 1, bs) if not motion else O.window_trajectories(X, t, **kw).astype(np.int64)
    idx = bf_patch_set(n, bs, step)
    assert np.array_equal(idx, O.patch_indices(n, bs, step))
    alpha, mu, sigma = 0.03, 0.02, 0.05  # as the driver holds them
    lams = [0.05, 0.3, 1.5, 8.0] if not exp_w else [0.5, 4.0, 30.0, 200.0]
    got = O.window_risk(X, t, [alpha, mu, sigma], lams, **kw)
    significant = {"fit": 0, "offset": 0, "div2": 0, "mean": 0}
    for lam, r in zip(lams, got):
        want, terms, N, sq = bf_risk_terms(u, traj, idx, bs, lam, exp_w, seed, alpha, mu, sigma)
        assert abs(r - want) <= 1e-10 * abs(want), (lam, r, want)
        # The test is only a pin if each term moves the risk by far more
        # than the tolerance: dropping a term, flipping its sign, or using
        # the unswapped (mu, sigma) must fail.
        assert terms["div1"] == 0.0  # U1 == U exactly (integer delta1 cube)
        for name in significant:
            significant[name] += abs(terms[name]) / N > 1e-6 * abs(want)
        unswapped = bf_risk_terms(u, traj, idx, bs, lam, exp_w, seed, alpha, sigma, mu)[0]
        assert abs(unswapped - want) > 1e-6 * abs(want)
    # the second difference vanishes where every rank is kept (small plain
    # lambda); it must matter at the other lambdas
    assert min(significant.values()) >= 2, significant


def test_risk_divergence_term_is_not_negligible():
    # The second-difference (divergence) term of src/pgure.hpp:136 changes
    # the risk at an intermediate lambda by far more than any tolerance used
    # for the oracle, so a dropped or mis-signed divergence fails the brute
    # force above.
    n, nf, bs = 8, 3, 2
    X = tiny_sequence(n, nf, 6)
    u = X / X.max()
    traj = motion_off_traj(n, 3, 1, bs)
    idx = bf_patch_set(n, bs, 1)
    want, terms, N, sq = bf_risk_terms(u, traj, idx, bs, 4.0, 1, 77, 0.03, 0.02, 0.05)
    assert abs(terms["div2"] / N) > 1e-5 * abs(want)  # the brute force agrees to 1e-10


# ----------------------------------------------------------------------------
# Thresholding closed form on a single-patch frame (src/svt.hpp:135-146,
# src/utils.hpp:96-106)
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("exp_w", [1, 0])
def test_threshold_closed_form_single_patch(exp_w):
    # A 4 x 4 frame holds exactly one 4 x 4 patch position, so the window is
    # one 16 x 15 Casorati matrix with known singular values and every
    # voxel's weight is 1: the reconstruction is U diag(s') V^T with
    #   exp: s' = max(s - s_max exp(-lambda s^2 / 2), 0)
    #   plain: s' = max(s - lambda, 0).
    # Catches exp(-lambda s^2) without the 1/2, s in place of s_max, a
    # threshold applied to s^2, and hard instead of soft thresholding.
    # The data must be non-negative: the leading singular vectors are the
    # constant vectors and s_1 / sqrt(16 * 15) exceeds the sum of the other
    # singular values, so every entry is positive.
    rng = np.random.default_rng(11)
    bs, nw = 4, 15
    Q1, _ = np.linalg.qr(np.column_stack([np.ones(16), rng.standard_normal((16, 14))]))
    Q2, _ = np.linalg.qr(np.column_stack([np.ones(15), rng.standard_normal((15, 14))]))
    s = np.array([400.0, 6.5, 4.0, 3.0, 2.5, 2.0, 1.6, 1.3, 1.0, 0.8, 0.6, 0.4, 0.3, 0.2, 0.1])
    A = Q1 @ np.diag(s) @ Q2.T
    assert A.min() > 0
    X = np.asfortranarray(np.stack([A[:, k].reshape(bs, bs, order="F") for k in range(nw)], axis=2))
    mx = X.max()
    su = s / mx  # singular values of the normalised window u = X / max(X)
    kw = dict(trajectory_length=15, patch_size=bs, motion_filter=-1, exponential_weighting=exp_w)
    lams = [50.0, 300.0, 1500.0, 8000.0] if exp_w else [0.005, 0.02, 0.1, 0.2]
    partial = 0
    for lam in lams:
        vol = O.window_reconstruct(X, 7, lam, **kw)
        if exp_w:
            sp = np.maximum(su - su.max() * np.exp(-0.5 * lam * su * su), 0.0)
        else:
            sp = np.maximum(su - lam, 0.0)
        B = Q1 @ np.diag(sp) @ Q2.T
        want = np.stack([B[:, k].reshape(bs, bs, order="F") for k in range(nw)], axis=2)
        assert np.abs(vol - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), lam
        partial += 1 < np.count_nonzero(sp) < nw  # some ranks below s_1 kept, some dropped
    assert partial >= 2


# ----------------------------------------------------------------------------
# ARPS hand traces (src/arps.hpp:145-375) on 3-frame windows, patch 2
# ----------------------------------------------------------------------------
# Window t = 1 of 3 frames: reference frame 1; the forward pass matches the
# blocks of frame 1 in frame 2 (motion slot 1), then the first backward pass
# matches the same blocks of frame 1 in frame 0 and reads the motion the
# forward pass left in slot 1 as its step predictor (src/arps.hpp:104-131).
# Source blocks are small values (0-10) and unplanted destination pixels
# large (50-100), so a candidate that includes any unplanted pixel costs far
# more than one that does not; every premise below is also asserted from the
# brute-force cost.

BS, N12, MW = 2, 12, 3


def _base(seed):
    rng = np.random.default_rng(seed)
    src = rng.uniform(0, 10, size=(N12, N12))
    d0 = rng.uniform(50, 100, size=(N12, N12))
    d2 = rng.uniform(50, 100, size=(N12, N12))
    return src, d0, d2


def _run(src, d0, d2):
    X = np.asfortranarray(np.stack([d0, src, d2], axis=2))
    tr = O.window_trajectories(X, 1, trajectory_length=3, patch_size=BS, motion_filter=-1, motion_estimation=1,
                               motion_window=MW)
    u = X / X.max()
    n1 = N12 - BS + 1

    def pos(k, i, j):
        return tuple(tr[k, i + n1 * j])

    def cost(k, i, j, y, x):
        return ((u[i:i + BS, j:j + BS, 1] - u[y:y + BS, x:x + BS, k]) ** 2).sum() / (BS * BS)

    return pos, cost


def _rood(y, x):
    return [(y - 1, x), (y, x - 1), (y, x + 1), (y + 1, x)]


def test_arps_first_column_uses_step_two():
    # Block (5, 0) (column 0): the large diamond has step 2 whatever the
    # predictor (src/arps.hpp:206-210).  Its exact copy sits 2 rows down in
    # frame 2 while its own position is a rood local minimum, so a one-step
    # search would stay; the step-2 point (7, 0) is taken.  The block (5, 4)
    # of the same pass (column > 0, zero predictor: step 0, the large diamond
    # skipped) with the same construction stays at (5, 4).
    src, d0, d2 = _base(1)
    for j in (0, 4):
        d2[7:9, j:j + 2] = src[5:7, j:j + 2]          # exact copy 2 rows down
        d2[5:7, j:j + 2] = src[5:7, j:j + 2] + 0.7    # own position: small cost
    pos, cost = _run(src, d0, d2)
    for j in (0, 4):
        c0 = cost(2, 5, j, 5, j)
        assert 0 < c0 and cost(2, 5, j, 7, j) == 0
        assert all(cost(2, 5, j, y, x) > c0 for y, x in _rood(5, j) if 0 <= x <= N12 - BS)
    assert all(cost(2, 5, 0, y, x) > 0 for y, x in _rood(7, 0) if x >= 0)
    assert pos(2, 5, 0) == (7, 0)
    assert pos(2, 5, 4) == (5, 4)


def test_arps_first_backward_pass_uses_forward_motion():
    # Block (5, 6): in frame 2 a descending rood path (5, 6) -> (5, 7) ->
    # (5, 8) leads to its exact copy, so the forward motion is (0, +2).  In
    # frame 0 the block's own position is a rood local minimum with a small
    # cost and its exact copy sits at (5, 4).  The first backward pass takes
    # step 2 from the forward motion (src/arps.hpp:211-228), so the large
    # diamond reaches (5, 4); without the predictor it would stay at (5, 6).
    src, d0, d2 = _base(2)
    src[5:7, 6:8] = [[1.0, 1.5], [2.0, 2.5]]
    d2[5:7, 8:10] = src[5:7, 6:8]      # exact copy at (5, 8)
    d2[5:7, 7] = src[5:7, 6]           # (5, 7) = [col 0, col 0]: near match
    d0[5:7, 4:6] = src[5:7, 6:8]       # exact copy at (5, 4)
    d0[5:7, 6:8] = src[5:7, 6:8] + 0.3  # own position: small cost
    pos, cost = _run(src, d0, d2)
    # forward premises: each move is the strict rood minimum of the path
    c = {p: cost(2, 5, 6, *p) for p in [(5, 6), (5, 7), (5, 8), (4, 6), (6, 6), (5, 5), (4, 7), (6, 7), (4, 8),
                                        (6, 8), (5, 9)]}
    assert c[(5, 7)] < min(c[(4, 6)], c[(6, 6)], c[(5, 5)], c[(5, 6)])
    assert c[(5, 8)] == 0 < min(c[(4, 7)], c[(6, 7)])
    assert 0 < min(c[(4, 8)], c[(6, 8)], c[(5, 9)])
    assert pos(2, 5, 6) == (5, 8)
    # backward premises: own position is a strict rood minimum, the copy is exact
    c0 = cost(0, 5, 6, 5, 6)
    assert 0 < c0 < min(cost(0, 5, 6, y, x) for y, x in _rood(5, 6))
    assert cost(0, 5, 6, 5, 4) == 0
    assert pos(0, 5, 6) == (5, 4)


def test_arps_large_diamond_tie_keeps_first_point():
    # Block (5, 0) in the backward pass (column 0, step 2): exact copies at
    # (3, 0) (point 0 of the diamond, row - step) and (7, 0) (point 4, row +
    # step) cost exactly 0 each; the first index wins (arma::find(costs ==
    # costs.min()) then point(0), src/arps.hpp:289-292), and the small diamond
    # around (3, 0) finds nothing cheaper.
    src, d0, d2 = _base(3)
    d0[3:5, 0:2] = src[5:7, 0:2]
    d0[7:9, 0:2] = src[5:7, 0:2]
    pos, cost = _run(src, d0, d2)
    assert cost(0, 5, 0, 3, 0) == 0 == cost(0, 5, 0, 7, 0)
    assert all(cost(0, 5, 0, y, x) > 0 for y, x in _rood(3, 0) if x >= 0)
    assert pos(0, 5, 0) == (3, 0)


def test_arps_small_diamond_tie_moves_to_earlier_point():
    # Block (8, 6), forward pass, zero predictor (step 0): rows 7-10 of
    # columns 6-7 in frame 2 hold one constant, so the block's own position,
    # the point above it (index 0) and the point below (index 4) cost
    # exactly the same.  The first index of the minimum is the point above,
    # so the search moves there (the centre has index 2) and then stops:
    # its own upper point costs more and the old centre is marked visited
    # (src/arps.hpp:299-368).
    src, d0, d2 = _base(4)
    d2[7:11, 6:8] = 5.0
    pos, cost = _run(src, d0, d2)
    c0 = cost(2, 8, 6, 8, 6)
    assert cost(2, 8, 6, 7, 6) == c0 == cost(2, 8, 6, 9, 6)
    assert min(cost(2, 8, 6, 8, 5), cost(2, 8, 6, 8, 7)) > c0
    assert min(cost(2, 8, 6, 6, 6), cost(2, 8, 6, 7, 5), cost(2, 8, 6, 7, 7)) > c0
    assert pos(2, 8, 6) == (7, 6)


# ----------------------------------------------------------------------------
# Noise estimator pieces (src/noise.hpp:173-383)
# ----------------------------------------------------------------------------

def test_wls_fit_recovers_exact_line():
    # weighted least squares with unit weights is exact on a line; the first
    # residual is then zero and the iteration stops (src/noise.hpp:328-383)
    x = np.linspace(0.01, 0.9, 57)
    a, b = O.noise_wls(x, 0.37 * x + 0.012)
    assert abs(a - 0.37) < 1e-12 and abs(b - 0.012) < 1e-12


def test_wls_fit_symmetric_residual_pairs():
    # Every x appears twice with residuals +-d(x) about a line.  The Huber
    # weights of a pair are equal, so every reweighted fit passes through the
    # pair midpoints: the line itself.  A sign error in the residuals or the
    # slope update moves it.
    x = np.repeat(np.linspace(0.05, 0.8, 20), 2)
    d = np.repeat(0.01 + 0.3 * np.linspace(0, 1, 20) ** 3, 2) * np.tile([1.0, -1.0], 20)
    a, b = O.noise_wls(x, 0.8 * x + 0.05 + d)
    assert abs(a - 0.8) < 1e-9 and abs(b - 0.05) < 1e-9


def test_robust_mean_closed_forms():
    # a constant is its own mean; a set symmetric about c keeps symmetric
    # weights in every iteration, so the weighted mean stays c
    assert abs(O.noise_robust_mean(np.full(37, 0.4321)) - 0.4321) <= 4 * np.spacing(0.4321)
    c = 0.25
    off = np.array([0.001, 0.004, 0.02, 0.05, 0.3, 1.7])
    assert abs(O.noise_robust_mean(np.concatenate([c + off, c - off, [c]])) - c) < 1e-12


def test_robust_var_closed_form():
    # (1.4826 * median |a - median(a)|)^2 (src/noise.hpp:232-236): for
    # 0..10 the median is 5 and the absolute deviations 5,4,...,0,...,5 have
    # median 3
    assert abs(O.noise_robust_var(np.arange(11.0)) - (1.4826 * 3.0) ** 2) < 1e-12


def test_compute_mode_histograms():
    # a = round(A * n / max(A)); the most frequent a wins, the first to reach
    # the count on ties; the result is a * max(A) / n (src/noise.hpp:273-301)
    # [5,5,5,1,9,10]: n = 6, scale 0.6 -> a = 3,3,3,1,5,6 -> mode 3 -> 3 * 10/6
    assert abs(O.noise_mode([5, 5, 5, 1, 9, 10]) - 5.0) < 1e-12
    # [2,8,8,2,10]: scale 0.5 -> 1,4,4,1,5: bins 1 and 4 tie at 2, bin 1 is
    # reached first -> 1 * 10/5
    assert abs(O.noise_mode([2, 8, 8, 2, 10]) - 2.0) < 1e-12
    # [8,2,8,2,10]: the same tie, bin 4 reached first -> 4 * 10/5
    assert abs(O.noise_mode([8, 2, 8, 2, 10]) - 8.0) < 1e-12


F0025 = {16: 1.27893, 32: 1.13046, 64: 1.06318}


def _stat(A):
    """F statistic of src/noise.hpp:186-217: data variance against the
    variance of the mirrored 5-point pseudo-residuals."""
    N = A.shape[0]
    idx = np.arange(N)
    p = np.where(idx + 1 == N, 1, idx + 1)
    m = np.where(idx == 0, N - 2, idx - 1)
    res = (5 * A - (A[p, :] + A[m, :] + A[:, m] + A[:, p])) / np.sqrt(30.0)
    Sz = ((A - A.mean()) ** 2).sum() / (N * N - 1)
    Se = ((res - res.mean()) ** 2).sum() / (N * N - 1)
    return Sz / Se if Sz > Se else Se / Sz


@pytest.mark.parametrize("N", [16, 32, 64])
def test_split_threshold_from_f_table(N):
    # A smooth ramp added to white noise raises the statistic continuously
    # from about 1; blocks just below and just above the 2.5 % F value of
    # the block's size split as the table dictates (a table looked up at the
    # wrong size, or a one-sided test, fails one of the two).
    rng = np.random.default_rng(N)
    E = rng.standard_normal((N, N))
    G = np.add.outer(np.arange(N), 2 * np.arange(N)) / N

    def block(t):
        return E + t * G

    lo, hi = 0.0, 50.0
    for _ in range(200):  # stat(t) is increasing: bisection for stat == F
        mid = 0.5 * (lo + hi)
        if _stat(block(mid)) > F0025[N]:
            hi = mid
        else:
            lo = mid
    below, above = block(lo * 0.999), block(hi * 1.001)
    assert _stat(below) < F0025[N] < _stat(above)
    assert not O.noise_split(below)
    assert O.noise_split(above)


def test_split_zero_block_and_minimum_size():
    # An all-zero block has zero data and residual variance: the statistic
    # is 0/0 and the comparison with the table value is false
    # (src/noise.hpp:214-220).  Blocks at the minimum size 8 never split,
    # whatever their statistic (src/noise.hpp:185-188).
    assert not O.noise_split(np.zeros((32, 32)))
    ramp = np.add.outer(np.arange(8.0), np.arange(8.0))
    assert _stat(ramp) > 1.64602 and not O.noise_split(ramp)  # 1.64602: the table's 8 x 8 value
    assert O.noise_split(np.add.outer(np.arange(16.0), np.arange(16.0)))


def test_quadtree_of_four_quadrants():
    # Four 16 x 16 quadrants of white noise on different offsets: the root
    # splits (its statistic is far above the 32 x 32 table value 1.13046),
    # and no quadrant does (each statistic is below the 16 x 16 value
    # 1.27893).  The recursion revisits the node that was last before the
    # split (here the root itself, which splits again and appends the four
    # quadrants a second time) and then the first three new nodes
    # (src/noise.hpp:433-455); the parents to shed are {0}, and node dele(0)
    # is never shed (src/noise.hpp:66-72).  So the root stays and every
    # quadrant appears twice.
    rng = np.random.default_rng(22)
    F = 0.01 * rng.standard_normal((32, 32))
    F[:16, :16] += 0.1
    F[16:, :16] += 0.2
    F[:16, 16:] += 0.3
    F[16:, 16:] += 0.4
    assert _stat(F) > F0025[32]
    q = [[0, 0, 16], [16, 0, 16], [0, 16, 16], [16, 16, 16]]
    for r, c, n in q:
        assert _stat(F[r:r + n, c:c + n]) < F0025[16]
    assert O.noise_tree(F).tolist() == [[0, 0, 32]] + q + q
    # a block that does not split is one node
    assert O.noise_tree(F[:16, :16]).tolist() == [[0, 0, 16]]


def test_noise_method1_too_few_leaves_is_an_error():
    # Method 1 reads the first round(0.05 * L) + 1 sorted leaf statistics, L =
    # pixels per frame / leaves.  An all-zero 256 x 256 window of 15 frames
    # has one leaf per frame (0/0 statistics never split): L = 4369, 219
    # statistics needed, 15 available.  The reference's bounds-checked copy
    # aborts (src/noise.hpp:115-121, 384-392); the oracle reports an error
    # instead of reading past the end.
    u = np.zeros((256, 256, 15))
    with pytest.raises(RuntimeError):
        O.noise_estimate(u, method=1)
