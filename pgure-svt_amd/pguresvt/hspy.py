"""HyperSpy adapter over SVT (pguresvt/hspy.py:7-80).

HyperSpy is optional: this module never imports it, it only calls the
signal's own methods, so any HyperSpy signal (or an object with the same
methods) can be passed.
"""

from .svt import SVT


class HSPYSVT(SVT):
    """SVT for HyperSpy signals of signal dimension 1 or 2.

    The data are handed to SVT.denoise as a spectrum stack (signal axis
    last), as the reference adapter does (pguresvt/hspy.py:22-56):

    * signal dimension 1: the axes-aligned data as they are;
    * signal dimension 2: the navigation space is unfolded, the stack is
      viewed as a Signal1D over navigation axis 0, its axes-aligned data are
      taken, and the signal is folded back.

    ``denoise`` returns a deep copy of the signal holding the denoised data,
    transposed back to the signal's order, titled "Denoised <title>"
    (pguresvt/hspy.py:58-80).
    """

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self._signal_type = None
        self._X = None

    def _prepare_to_denoise(self, signal):
        ndim = signal.axes_manager.signal_dimension
        if ndim == 1:
            self._signal_type = "spectrum"
            self._X = signal._data_aligned_with_axes
        elif ndim == 2:
            self._signal_type = "image"
            signal.unfold_navigation_space()
            try:
                self._X = signal.as_signal1D(spectral_axis=0)._data_aligned_with_axes
            finally:
                signal.fold()
        else:
            raise NotImplementedError(f"Expected 1D or 2D signal - got dimension {ndim}")

    def denoise(self, signal):
        self._prepare_to_denoise(signal)
        super().denoise(self._X)
        back = (1, 0) if self._signal_type == "spectrum" else (2, 0, 1)
        out = signal._deepcopy_with_new_data(self.Y_.transpose(back))
        out.metadata.General.title = f"Denoised {signal.metadata.General.title}".strip()
        return out
