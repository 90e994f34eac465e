"""+f_numbers/+distance: F-numbers keeping a fixed angular distance to the grating lobe."""
from __future__ import annotations

import math

from . import f_number


class constant(f_number):
    """+f_numbers/+distance/constant.m (distance_deg, F_number_ub = inf)."""
    kind = "distance.constant"

    def __init__(self, distance_deg: float = 5.0, F_number_ub: float = math.inf):
        if not (0 <= distance_deg <= 90) or not (F_number_ub > 0):
            raise ValueError("invalid distance F-number parameters (constant.m:20-21)")
        self.distance_deg = float(distance_deg)
        self.F_number_ub = float(F_number_ub)

    def _params(self):
        return (self.distance_deg, self.F_number_ub, 0.0)

    def __str__(self):
        return f"distance_deg_{self.distance_deg:.2f}_F_ub_{self.F_number_ub:.2f}"


class none(constant):
    """+f_numbers/+distance/none.m: distance.constant(0, inf)."""
    kind = "distance.none"

    def __init__(self):
        super().__init__(0.0, math.inf)
