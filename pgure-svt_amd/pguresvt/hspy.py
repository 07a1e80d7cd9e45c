"""HyperSpy adapter: denoise a signal's data with SVT (pguresvt/hspy.py:7-80).

HyperSpy is optional; this module imports it lazily.
"""

import numpy as np

from .svt import SVT


class HSPYSVT(SVT):
    """SVT for HyperSpy Signal2D image stacks (navigation axis = time).

    ``denoise(signal)`` runs SVT on the (x, y, t) view of ``signal.data``
    (shape (t, y, x)) and returns a new signal with the denoised data; plain
    numpy arrays of shape (t, y, x) are accepted too.
    """

    def denoise(self, signal):
        data = getattr(signal, "data", signal)
        arr = np.asarray(data)
        if arr.ndim != 3:
            raise ValueError("expected a stack of images with shape (t, y, x)")
        super().denoise(np.transpose(arr, (2, 1, 0)))
        out = np.transpose(self.Y_, (2, 1, 0))
        if hasattr(signal, "_deepcopy_with_new_data"):
            return signal._deepcopy_with_new_data(out)
        return out
