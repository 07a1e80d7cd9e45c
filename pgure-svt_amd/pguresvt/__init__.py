"""pguresvt on AMD MI355X: patchwise singular value thresholding with the
PGURE risk, as hand-written gfx950 kernels behind a C ABI (libpgsvt.so)."""

from .svt import SVT, mixed_noise_model

__all__ = ["mixed_noise_model", "SVT"]
