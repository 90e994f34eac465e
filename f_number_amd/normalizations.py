"""Normalizations of the apodization weights (mirror of +normalizations)."""
from __future__ import annotations

from . import _lib
from .windows import boxcar, window


class normalization:
    """Abstract normalization (+normalizations/normalization.m)."""

    code: int = 1
    window: window = boxcar()

    def __str__(self) -> str:
        return type(self).__name__


class off(normalization):
    """+normalizations/off.m: weights used as they are."""
    code = _lib.NORMALIZATION["off"]


class on(normalization):
    """+normalizations/on.m: weights divided by their sum over the elements."""
    code = _lib.NORMALIZATION["on"]


class windowed(on):
    """+normalizations/windowed.m: 'on' followed by a taper over the valid frequencies."""
    code = _lib.NORMALIZATION["windowed"]

    def __init__(self, w: window):
        if not isinstance(w, window):
            raise _lib.FNumberError("windowed:NoWindow", "window must be windows.window!")
        self.window = w
