"""Error paths forced from the host (PGSVT_FAULT): a device allocation or a
HIP call that fails returns PGSVT_E_HIP with a message and leaves the
library usable; an allocation that fails while a batch of windows is
allocated halves the batch and the call still succeeds with the same
frames.  Every kernel runs to completion: the failures are reported by the
host after the real calls succeeded."""

import numpy as np
import pytest

import pgsvt_lib as G
from pguresvt import _pguresvt as P
from pguresvt import phantom

pytestmark = pytest.mark.gpu

KW = dict(optimize_pgure=False, lambda1=2.0, motion_filter=-1)


@pytest.fixture(scope="module")
def seq():
    return phantom.noisy_sequence(48, 16, drift=1, alpha=1.0, sigma=6.0, seed=31)[1]


def test_injected_failures(gpu, seq, monkeypatch):
    Y0, e0 = G.denoise(seq, **KW)
    # the first buffer of the call (the input's device copy)
    monkeypatch.setenv("PGSVT_FAULT", "alloc:1")
    with pytest.raises(P.PGSVTError, match="error -2.*injected"):
        G.denoise(seq, **KW)
    # a HIP call in the middle of the decomposition
    monkeypatch.setenv("PGSVT_FAULT", "hip:12")
    with pytest.raises(P.PGSVTError, match="error -2.*injected"):
        G.denoise(seq, **KW)
    # PGURE search with the failure late in the search rounds
    monkeypatch.setenv("PGSVT_FAULT", "hip:60")
    with pytest.raises(P.PGSVTError, match="error -2"):
        G.denoise(seq, motion_filter=-1)
    # the eighth buffer is one of the batch's: the batch is halved and the
    # call succeeds with the frames of a clean run
    monkeypatch.setenv("PGSVT_FAULT", "alloc:8")
    Y1, e1 = G.denoise(seq, **KW)
    assert np.array_equal(Y1, Y0) and np.array_equal(e1, e0)
    # and with no fault the library works as before
    monkeypatch.delenv("PGSVT_FAULT")
    Y2, e2 = G.denoise(seq, **KW)
    assert np.array_equal(Y2, Y0) and np.array_equal(e2, e0)
    Yp, ep = G.denoise(seq, motion_filter=-1)
    assert np.isfinite(Yp).all() and (ep[:, 0] > 0).all()
