"""+f_numbers/+directivity: F-numbers from the element directivity."""
from __future__ import annotations

from . import f_number


class directivity(f_number):
    """+f_numbers/+directivity/directivity.m: F = 1 / (2 tan alpha), alpha where the
    element directivity drops by attenuation_dB (fminbnd on [0, pi/2])."""

    obliquity = "soft"

    def __init__(self, width_over_pitch: float = 0.918, attenuation_dB: float = 3.0):
        if not (0 <= width_over_pitch <= 1) or not (attenuation_dB > 0):
            raise ValueError("invalid directivity parameters (directivity.m:16-18)")
        self.width_over_pitch = float(width_over_pitch)
        self.attenuation_dB = float(attenuation_dB)
        self.thresh = 10 ** (-self.attenuation_dB / 20)

    def _params(self):
        return (self.width_over_pitch, self.attenuation_dB, 0.0)

    def __str__(self):
        return f"directivity_{self.obliquity}_fill_{self.width_over_pitch * 1e2:.2f}_att_{self.attenuation_dB:.1f}_dB"


class soft(directivity):
    """+f_numbers/+directivity/soft.m: obliquity cos(theta)."""
    kind = "directivity.soft"
    obliquity = "soft"


class rigid(directivity):
    """+f_numbers/+directivity/rigid.m: obliquity 1."""
    kind = "directivity.rigid"
    obliquity = "rigid"


class perrot(soft):
    """+f_numbers/+directivity/perrot.m: soft baffle, 3 dB."""
    kind = "directivity.perrot"

    def __init__(self, width_over_pitch: float):
        super().__init__(width_over_pitch, 3.0)


class szabo(rigid):
    """+f_numbers/+directivity/szabo.m: F = width_over_pitch * r / 1.2."""
    kind = "directivity.szabo"

    def __init__(self, width_over_pitch: float):
        super().__init__(width_over_pitch, 6.0)
