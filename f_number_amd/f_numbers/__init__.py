"""F-numbers (mirror of the reference's +f_numbers package).

Values come from the C++ port in libfnumber_hip.so (fn_f_number_values),
which replaces the MEX's mexCallMATLAB('compute_values') round trip
(gpu_bf_helper_functions.cu:35-51).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib


class f_number:
    """Abstract F-number (+f_numbers/f_number.m)."""

    kind: str = ""

    def _params(self):
        return (0.0, 0.0, 0.0)

    def _c(self) -> _lib.FNumber:
        p = self._params()
        return _lib.FNumber(_lib.F_NUMBER[self.kind], *[float(v) for v in p])

    def compute_values(self, element_pitch_over_lambda) -> np.ndarray:
        """F-number per pitch-to-wavelength ratio (f_number.m:50-114)."""
        r = np.ascontiguousarray(element_pitch_over_lambda, dtype=np.float64)
        if r.size == 0:
            raise _lib.FNumberError("compute_values:NonEmpty", "element_pitch_over_lambda must be nonempty")
        out = np.empty_like(r)
        F = self._c()
        _lib.check(_lib.load().fn_f_number_values(C.byref(F), r.ctypes.data_as(C.POINTER(C.c_double)), r.size,
                                                  out.ctypes.data_as(C.POINTER(C.c_double))), "compute_values")
        return out

    # f_number.m:119-182: main-lobe bound 1/sqrt(1 + (2F)^2)
    def compute_bounds_main(self, element_pitch_over_lambda) -> np.ndarray:
        values = self.compute_values(element_pitch_over_lambda)
        return 1 / np.sqrt(1 + (2 * values) ** 2)

    # f_number.m:187-240: distance of the first grating lobe, 1/r - bound
    def compute_distances_grating(self, element_pitch_over_lambda) -> np.ndarray:
        r = np.asarray(element_pitch_over_lambda, dtype=np.float64)
        return 1 / r - self.compute_bounds_main(r)

    # f_number.m:245-298: depth of field over element pitch, 6.1434 F^2 / r
    def compute_depths_of_field(self, element_pitch_over_lambda) -> np.ndarray:
        r = np.asarray(element_pitch_over_lambda, dtype=np.float64)
        return 6.1434 / r * self.compute_values(r) ** 2


class constant(f_number):
    """+f_numbers/constant.m: F = value for every frequency."""
    kind = "constant"

    def __init__(self, value: float = 1.0):
        if not (value >= 0):
            raise ValueError("value must be nonnegative")
        self.value = float(value)

    def _params(self):
        return (self.value, 0.0, 0.0)

    def __str__(self):
        return f"constant_{self.value:.2f}"


from . import directivity, distance, grating  # noqa: E402,F401
