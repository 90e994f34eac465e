"""Compute platforms (mirror of +platforms).

Only GPUs are served: the reference's CPU branch (das_pw.m:302-485,
das_sa.m:289-509) is not part of this package.
"""
from __future__ import annotations


class platform:
    """Abstract platform (+platforms/platform.m)."""


class cpu(platform):
    """+platforms/cpu.m.  Accepted for argument compatibility; das_pw/das_sa raise for it."""

    def __str__(self) -> str:
        return "cpu"


class gpu(platform):
    """+platforms/gpu.m: GPU selected by index."""

    def __init__(self, index: int = 0):
        index = int(index)
        if index < 0:
            raise ValueError("index must be nonnegative")  # gpu.m:16 mustBeNonnegative
        self.index = index

    def __str__(self) -> str:
        return f"gpu_{self.index}"
