"""Receive/transmit apodization windows (mirror of the reference's +windows package).

Samples are evaluated by the C++ port in libfnumber_hip.so
(fn_window_samples), with the semantics of +windows/<name>.m: boxcar and
tukey are truncated at |t| >= 1, hann and triangular are not (window.m:48-93).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class window:
    """Abstract window (+windows/window.m)."""

    id: int = -1
    parameter: float = 0.0

    def compute_samples(self, positions_over_halfwidth, semantics: str = "cpu") -> np.ndarray:
        """Samples of |positions_over_halfwidth| (window.m:48-93)."""
        t = np.ascontiguousarray(positions_over_halfwidth, dtype=np.float64)
        out = np.empty_like(t)
        w = self._c()
        L = _lib.load()
        _lib.check(L.fn_window_samples(C.byref(w), _lib.SEMANTICS[semantics],
                                       t.ctypes.data_as(C.POINTER(C.c_double)), t.size,
                                       out.ctypes.data_as(C.POINTER(C.c_double))), "compute_samples")
        return out

    def _c(self) -> _lib.Window:
        return _lib.Window(self.id, float(self.parameter))

    def __str__(self) -> str:
        return type(self).__name__


class boxcar(window):
    """+windows/boxcar.m: 1 for |t| < 1."""
    id = 0


class hann(window):
    """+windows/hann.m: (1 + cos(pi |t|)) / 2."""
    id = 1


class triangular(window):
    """+windows/triangular.m: 1 - |t|."""
    id = 4


class tukey(window):
    """+windows/tukey.m: flat for |t| <= 1 - fraction_cosine, cosine taper to |t| = 1."""
    id = 2

    def __init__(self, fraction_cosine: float = 0.5):
        fraction_cosine = float(fraction_cosine)
        # tukey.m:46-57
        if fraction_cosine <= 0:
            raise _lib.FNumberError("tukey:InvalidFractionsCosine", "Use windows.boxcar for a cosine fraction of zero!")
        if fraction_cosine >= 1:
            raise _lib.FNumberError("tukey:InvalidFractionsCosine", "Use windows.hann for a cosine fraction of one!")
        self.fraction_cosine = fraction_cosine
        self.fraction_rectangle = 1 - fraction_cosine
        self.parameter = fraction_cosine

    def __str__(self) -> str:
        return f"tukey_{self.fraction_cosine * 1e2:.2f}"


def compute_samples(w: window, positions_over_halfwidth, semantics: str = "cpu") -> np.ndarray:
    return w.compute_samples(positions_over_halfwidth, semantics)
