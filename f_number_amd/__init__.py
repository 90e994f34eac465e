"""f_number_amd: MI355X-native Fourier-domain delay-and-sum beamforming with a
frequency-dependent F-number (the das_pw / das_sa path of mschiffn/f_number).

The computation runs in libfnumber_hip.so (hand-written HIP for gfx950);
this package mirrors the reference's MATLAB interface on top of its C ABI.
"""
from . import f_numbers, normalizations, platforms, windows  # noqa: F401
from ._lib import FNumberError, LIB_PATH, load  # noqa: F401
from .das import das_pw, das_sa  # noqa: F401
from .plan import Batch, PwPlan, SaPlan  # noqa: F401
from . import autofocus, examples, synth, tuning  # noqa: F401,E402
from .linescan import LineScan, das_lbl  # noqa: F401,E402

__all__ = ["das_pw", "das_sa", "PwPlan", "SaPlan", "windows", "f_numbers", "normalizations", "platforms",
           "FNumberError", "load", "Batch", "LineScan", "das_lbl"]
