"""SVT estimator and Poisson-Gaussian noise model.

Public interface of the reference package (pguresvt/svt.py:14-337): the same
class, constructor keywords and defaults, fitted attributes and validation
messages, running on the MI355X library through the pguresvt_* entry points.
"""

import numpy as np

from ._pguresvt import pguresvt_d, pguresvt_f, pguresvt_u8, pguresvt_u16


def _is_power_of_two(n):
    return n != 0 and (n & (n - 1)) == 0


def mixed_noise_model(X, alpha=1.0, mu=0.0, sigma=0.0, random_state=None):
    """Add Poisson-Gaussian noise: Y = alpha * Poisson(X0 / alpha) + N(mu, sigma^2).

    X is first scaled to a maximum of 1 and the result scaled back
    (pguresvt/svt.py:14-74).  ``alpha`` must lie in (0, 1], ``sigma`` >= 0.
    ``random_state`` is None, an int seed or a numpy RandomState; draws are
    the Poisson field followed by the normal field.
    """
    if alpha <= 0.0 or alpha > 1.0:
        raise ValueError("alpha should be in range [0, 1]")
    if sigma < 0.0:
        raise ValueError("sigma should be >= 0.0")
    rng = random_state if isinstance(random_state, np.random.RandomState) else np.random.RandomState(random_state)
    Xf = X.astype(float)
    scale = Xf.max()
    Xf /= scale
    Y = alpha * rng.poisson(Xf / alpha) + mu + sigma * rng.normal(size=Xf.shape)
    Y *= scale
    return Y


class SVT:
    """Singular value thresholding for denoising image sequences (PGURE-SVT).

    Parameters and attributes follow the reference estimator
    (pguresvt/svt.py:77-191): trajectory_length, patch_size, patch_overlap,
    motion_estimation, motion_window, motion_filter, optimize_pgure, lambda1,
    exponential_weighting, noise_method, noise_alpha, noise_mu, noise_sigma,
    tol, max_iter, n_jobs, random_seed.  After ``denoise(X)``: ``Y_`` (the
    denoised sequence, same shape as X), ``lambda1s_``, ``noise_alphas_``,
    ``noise_mus_``, ``noise_sigmas_`` (one value per frame).
    """

    def __init__(self, trajectory_length=15, patch_size=4, patch_overlap=1, motion_estimation=True,
                 motion_window=7, motion_filter=5, optimize_pgure=True, lambda1=None, exponential_weighting=True,
                 noise_method=4, noise_alpha=None, noise_mu=None, noise_sigma=None, tol=1e-7, max_iter=500,
                 n_jobs=None, random_seed=None):
        self.trajectory_length = trajectory_length
        self.patch_size = patch_size
        self.patch_overlap = patch_overlap
        self.motion_estimation = motion_estimation
        self.motion_window = motion_window
        self.motion_filter = motion_filter
        self.optimize_pgure = optimize_pgure
        self.lambda1 = lambda1
        self.exponential_weighting = exponential_weighting
        self.noise_method = noise_method
        self.noise_alpha = noise_alpha
        self.noise_mu = noise_mu
        self.noise_sigma = noise_sigma
        self.tol = tol
        self.max_iter = max_iter
        self.n_jobs = n_jobs
        self.random_seed = random_seed
        self.Y_ = None

    def _check_arguments(self, X):
        """Validate arguments exactly as the reference does before calling C++."""
        if X.min() < 0.0:
            raise ValueError(
                "Negative values found in data. PGURE-SVT requires strictly non-negative image data."
            )

        def sentinel(v):
            return -1 if v is None else v

        self.lambda1_ = sentinel(self.lambda1)
        self.motion_filter_ = sentinel(self.motion_filter)
        self.noise_alpha_ = sentinel(self.noise_alpha)
        self.noise_mu_ = sentinel(self.noise_mu)
        self.noise_sigma_ = sentinel(self.noise_sigma)
        self.n_jobs_ = sentinel(self.n_jobs)
        self.random_seed_ = sentinel(self.random_seed)

        if self.patch_overlap > self.patch_size:
            raise ValueError(
                f"Invalid patch_overlap parameter: got {self.patch_overlap}, "
                f"should not be greater than patch_size ({self.patch_size})"
            )
        if self.trajectory_length % 2 == 0 or self.trajectory_length < 1:
            raise ValueError(
                f"Invalid trajectory_length parameter: got {self.trajectory_length},"
                "but expected a positive, odd-valued integer"
            )
        if self.motion_estimation:
            if self.motion_window < 2 or self.motion_window % 2 == 0:
                raise ValueError(
                    f"Invalid motion_window parameter: got {self.motion_window}, "
                    "should be greater a positive, odd-valued integer > 1 pixel."
                )
            if not isinstance(self.motion_filter_, int):
                raise ValueError(
                    f"Invalid motion_filter parameter: got {type(self.motion_filter)}, "
                    "should be an integer number of pixels or None."
                )
        if not self.optimize_pgure and (self.lambda1 is None or self.lambda1 < 0.0):
            raise ValueError(
                f"Invalid lambda1 parameter: got {self.lambda1}, "
                "should be a float >= 0.0 if optimize_pgure is None."
            )
        if any(v < 0.0 for v in (self.noise_alpha_, self.noise_mu_, self.noise_sigma_)):
            if X.shape[0] != X.shape[1]:
                raise ValueError(f"Quadtree noise estimation requires square images, got {X.shape}")
            if not _is_power_of_two(X.shape[0]):
                raise ValueError("Quadtree noise estimation requires image dimensions 2^N")

    def denoise(self, X):
        """Denoise X (shape (nx, ny, nt); uint8, uint16, float32 or float64).

        Returns the estimator with ``Y_`` and the per-frame estimates set.
        """
        self._check_arguments(X)
        entry = {
            np.dtype("uint8"): pguresvt_u8,
            np.dtype("uint16"): pguresvt_u16,
            np.dtype("float32"): pguresvt_f,
            np.dtype("float64"): pguresvt_d,
        }
        dt = getattr(X, "dtype", None)
        if dt not in entry:
            raise TypeError(f"Invalid dtype: got {dt}, but only {list(entry.keys())} are supported")
        if not X.flags.f_contiguous:
            X = np.asfortranarray(X, dtype=dt)
        Y, est, _ = entry[dt](
            input_images=X,
            trajectory_length=self.trajectory_length,
            patch_size=self.patch_size,
            patch_overlap=self.patch_overlap,
            motion_estimation=self.motion_estimation,
            motion_window=self.motion_window,
            motion_filter=self.motion_filter_,
            optimize_pgure=self.optimize_pgure,
            lambda1=self.lambda1_,
            exponential_weighting=self.exponential_weighting,
            noise_method=self.noise_method,
            noise_alpha=self.noise_alpha_,
            noise_mu=self.noise_mu_,
            noise_sigma=self.noise_sigma_,
            tol=self.tol,
            max_iter=self.max_iter,
            n_jobs=self.n_jobs_,
            random_seed=self.random_seed_,
        )
        # back to the input orientation, C-contiguous
        self.Y_ = np.ascontiguousarray(np.transpose(Y, (2, 1, 0)))
        self.lambda1s_ = est[0]
        self.noise_alphas_ = est[1]
        self.noise_mus_ = est[2]
        self.noise_sigmas_ = est[3]
        return self
