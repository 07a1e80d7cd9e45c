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


# ------------------------------------------------------- windows and model --

class _Model(C.Structure):
    """Mirror of pgsvt_lambda_model (include/pgsvt.h)."""

    _fields_ = [("n_in", C.c_uint32), ("hidden", C.c_uint32), ("mean", C.c_void_p), ("inv_std", C.c_void_p),
                ("params", C.c_void_p), ("y_mean", C.c_float), ("y_std", C.c_float)]


def _bind_pipeline(L):
    if not getattr(L, "_learned_pipeline_bound", False):
        vp, u32 = C.c_void_p, C.c_uint32
        P = C.POINTER(_P.pgsvt_params)
        L.pgsvt_denoise_range.argtypes = [vp, C.c_int, u32, u32, u32, P, u32, u32, vp, vp]
        L.pgsvt_learned_features.argtypes = [vp, C.c_int, u32, u32, u32, P, u32, u32, vp, vp]
        L.pgsvt_denoise_learned.argtypes = [vp, C.c_int, u32, u32, u32, P, C.POINTER(_Model), vp, vp]
        L.pgsvt_set_learned_fallback.argtypes = [C.c_double, C.c_double, C.c_uint32]
        L.pgsvt_learned_fallbacks.argtypes = []
        L._learned_pipeline_bound = True
    return L


def _cube(X):
    X = np.asfortranarray(X)
    dt = {np.dtype("uint8"): 0, np.dtype("uint16"): 1, np.dtype("float32"): 2, np.dtype("float64"): 3}[X.dtype]
    return X, dt


def window_length(patch_size, trajectory_length):
    """Frames per window: Nt clamped to bs^2 - 1, then 2 floor(Nt / 2) + 1
    (src/pguresvt.hpp:57-58)."""
    bs2 = patch_size * patch_size
    nt = bs2 - 1 if bs2 < trajectory_length else trajectory_length
    return 2 * (nt // 2) + 1


def window_risk(X, t, noise, lambdas, **params):
    """PGURE risk of window t at each lambda (pgsvt_window_risk); noise =
    (alpha, mu, sigma) as the driver estimated them."""
    X, dt = _cube(X)
    nr, nc, nf = X.shape
    L = _bind_pipeline(_lib())
    p = _P.make_params(**params)
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    nz = np.ascontiguousarray(noise, dtype=np.float64)
    out = np.zeros(lam.size)
    _P.check(L.pgsvt_window_risk(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t, nz.ctypes.data, lam.ctypes.data,
                                 lam.size, out.ctypes.data))
    return out


def n_features(window):
    """Feature count for windows of `window` frames."""
    return N_SAMPLE * window + 3


def denoise_range(X, t_begin, t_end, **params):
    """The reference driver over output frames t_begin..t_end-1 of X
    (rows, cols, frames): (Y (rows, cols, t_end - t_begin), estimates (count, 4))."""
    X, dt = _cube(X)
    nr, nc, nf = X.shape
    L = _bind_pipeline(_lib())
    p = _P.make_params(**params)
    Y = np.zeros((nr, nc, t_end - t_begin), order="F")
    est = np.zeros((t_end - t_begin, 4), order="F")
    _P.check(L.pgsvt_denoise_range(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t_begin, t_end, Y.ctypes.data,
                                   est.ctypes.data))
    return Y, est


def features(X, t_begin, t_end, **params):
    """Raw learned-lambda features of output frames t_begin..t_end-1:
    (features (count, n_in) float32, estimates (count, 4))."""
    X, dt = _cube(X)
    nr, nc, nf = X.shape
    L = _bind_pipeline(_lib())
    p = _P.make_params(**params)
    window = window_length(p.patch_size, p.trajectory_length)
    F = np.zeros((t_end - t_begin, n_features(window)), dtype=np.float32)
    est = np.zeros((t_end - t_begin, 4), order="F")
    _P.check(L.pgsvt_learned_features(X.ctypes.data, dt, nr, nc, nf, C.byref(p), t_begin, t_end, F.ctypes.data,
                                      est.ctypes.data))
    return F, est


class LambdaModel:
    """Trained proposal network with its feature and target standardisation."""

    def __init__(self, params, mean, std, y_mean, y_std, hidden=HIDDEN):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.mean = np.ascontiguousarray(mean, dtype=np.float32)
        self.std = np.ascontiguousarray(std, dtype=np.float32)
        self.inv_std = np.ascontiguousarray(1.0 / self.std, dtype=np.float32)
        self.y_mean = float(y_mean)
        self.y_std = float(y_std)
        self.hidden = int(hidden)
        self.n_in = self.mean.size
        assert self.params.size == param_count(self.n_in, self.hidden)

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, std=self.std, y=np.array([self.y_mean, self.y_std]),
                 hidden=np.array([self.hidden]))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], z["mean"], z["std"], z["y"][0], z["y"][1], int(z["hidden"][0]))

    def _c(self):
        return _Model(self.n_in, self.hidden, self.mean.ctypes.data, self.inv_std.ctypes.data,
                      self.params.ctypes.data, self.y_mean, self.y_std)

    def predict_log_lambda(self, F):
        """log(lambda) for raw feature rows F (n, n_in)."""
        Z = (np.asarray(F, dtype=np.float32) - self.mean) * self.inv_std
        return self.y_mean + self.y_std * mlp_forward(self.params, Z, self.n_in, self.hidden)


def default_model_path():
    import os

    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "lambda_model.npz")


FALLBACK_DELTA, FALLBACK_GAP, FALLBACK_REFINE = 0.1, 0.01, 2  # the library's defaults


def set_fallback(delta=FALLBACK_DELTA, gap=FALLBACK_GAP, refine=FALLBACK_REFINE):
    """The learned path's probe (pgsvt_set_learned_fallback): risks at
    lambda / (1 + delta), lambda and lambda (1 + delta); a window whose
    predicted relative risk gap exceeds `gap` moves to the probe's estimate
    of the minimum and is probed again, up to `refine` times, then
    searched.  gap <= 0: no probe, every proposal kept.  Process-wide."""
    L = _bind_pipeline(_lib())
    _P.check(L.pgsvt_set_learned_fallback(float(delta), float(gap), int(refine)))


def last_fallbacks():
    """Windows searched by this thread's last denoise_learned call."""
    return int(_bind_pipeline(_lib()).pgsvt_learned_fallbacks())


def denoise_learned(X, model=None, **params):
    """Denoise X (rows, cols, frames) with learned lambdas: (Y, estimates
    (frames, 4): lambda (the proposal, or the search's where the probe sent
    the window to the search), alpha, mu, sigma)."""
    if model is None:
        model = LambdaModel.load(default_model_path())
    X, dt = _cube(X)
    nr, nc, nf = X.shape
    L = _bind_pipeline(_lib())
    p = _P.make_params(**params)
    Y = np.zeros((nr, nc, nf), order="F")
    est = np.zeros((nf, 4), order="F")
    m = model._c()
    _P.check(L.pgsvt_denoise_learned(X.ctypes.data, dt, nr, nc, nf, C.byref(p), C.byref(m), Y.ctypes.data,
                                     est.ctypes.data))
    return Y, est


def fit(F, log_lambda, epochs=150, batch=256, lr=1e-3, seed=0, hidden=HIDDEN):
    """Standardise features and targets, train on the GPU; returns
    (LambdaModel, loss per epoch)."""
    F = np.asarray(F, dtype=np.float32)
    y = np.asarray(log_lambda, dtype=np.float64)
    mean = F.mean(0)
    std = F.std(0)
    std[std < 1e-6] = 1.0
    y_mean, y_std = float(y.mean()), float(y.std() or 1.0)
    Z = (F - mean) / std
    t = ((y - y_mean) / y_std).astype(np.float32)
    p, losses = mlp_train(Z, t, init_params(F.shape[1], hidden, seed), epochs, batch, lr, seed, hidden)
    return LambdaModel(p, mean, std, y_mean, y_std, hidden), losses


# --------------------------------------------------------------- training data --

TRAIN_PARAMS = dict(trajectory_length=15, patch_size=4, patch_overlap=1, motion_window=7, motion_filter=-1,
                    noise_method=4, max_iter=500, n_jobs=-1, random_seed=101, optimize_pgure=1,
                    exponential_weighting=1, motion_estimation=1, lambda1=-1.0, noise_alpha=-1.0, noise_mu=-1.0,
                    noise_sigma=-1.0, tol=1e-7)


def training_sequence(seed, n=128, frames=30):
    """Seeded synthetic sequence of the training distribution: blob phantom
    (4-16 blobs, drift +-2 px per frame), Poisson-Gaussian noise with
    alpha in [0.1, 4], sigma in [1, 20], mu = 0."""
    from . import phantom

    rng = np.random.default_rng(seed)
    blobs = int(rng.integers(4, 17))
    alpha = float(rng.uniform(0.1, 4.0))
    sigma = float(rng.uniform(1.0, 20.0))
    _, noisy = phantom.noisy_sequence(n, frames, blobs=blobs, drift=2, alpha=alpha, mu=0.0, sigma=sigma, seed=seed)
    return noisy


def label_window(X, t, **params):
    """(raw features, PGURE-optimal lambda from the GPU search) of window t."""
    kw = dict(TRAIN_PARAMS)
    kw.update(params)
    _, est = denoise_range(X, t, t + 1, **kw)
    F, _ = features(X, t, t + 1, **kw)
    return F[0], est[0, 0]
