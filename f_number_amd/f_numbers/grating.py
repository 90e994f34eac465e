"""+f_numbers/+grating: F-numbers that bound the first grating lobe."""
from __future__ import annotations

import math

from . import f_number


class grating(f_number):
    """+f_numbers/+grating/grating.m: chi (deg), F_ub, anti-aliasing distance delta (deg)."""

    def __init__(self, angle_lb_deg: float, F_number_ub: float, distance_deg: float = 10.0):
        if not (0 <= angle_lb_deg <= 90) or not (F_number_ub >= 0) or not (distance_deg <= 90):
            raise ValueError("invalid grating F-number parameters (grating.m:16-18)")
        self.angle_lb_deg = float(angle_lb_deg)
        self.F_number_ub = float(F_number_ub)
        self.distance_deg = float(distance_deg)
        # grating.m:72-75
        self.thresh = math.sin(math.pi / 180 * self.angle_lb_deg)
        self.lower_bound = 1 / (1 + self.thresh)
        self.upper_bound = 1 / (1 / math.sqrt(1 + (2 * self.F_number_ub) ** 2) + self.thresh)
        self.critical = 1 / self.thresh if self.thresh else math.inf

    def _params(self):
        return (self.angle_lb_deg, self.F_number_ub, self.distance_deg)


class angle_lb(grating):
    """+f_numbers/+grating/angle_lb.m: the proposed frequency-dependent F-number."""
    kind = "grating.angle_lb"

    def __str__(self):
        s = f"angle_lb_deg_{self.angle_lb_deg:.2f}_F_ub_{self.F_number_ub:.2f}"
        if self.distance_deg >= 0:
            s += f"_distance_deg_{self.distance_deg:.2f}"
        return s
