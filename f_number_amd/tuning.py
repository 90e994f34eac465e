"""Tiling cache of the autotuner: one measurement per plan shape.

PwPlan/SaPlan.autotune times the candidate tilings (transmit tiles per
workgroup, bin split) of one plan; plans of the same shape (grid, elements,
transmit rows, bins, precision, window class) share the result.  The cache can
be saved to / loaded from a JSON file so a later process starts tuned.
"""
from __future__ import annotations

import json
import os


def key(plan) -> str:
    """Shape signature of a plan (what the tiling choice depends on)."""
    inf = plan.info()
    return "|".join(str(v) for v in (type(plan).__name__, inf["n_voxels"], inf["n_tx"], inf["n_f"], inf["k_steps"],
                                    getattr(plan, "n_frames", 1), getattr(plan, "precision", "fp32"),
                                    getattr(plan, "window_class", "")))


class TilingCache:
    def __init__(self, path: str | None = None):
        self.path = path
        self.table: dict[str, dict] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.table = json.load(f)

    def tune(self, plan, rf, image, repetitions: int = 1) -> dict:
        """Apply the cached tiling of the plan's shape, or autotune and remember it."""
        k = key(plan)
        hit = self.table.get(k)
        if hit is not None:
            plan.set_tiling(hit["m_tiles"], hit["bin_split"])
            return dict(hit, cached=True)
        res = plan.autotune(rf, image, repetitions=repetitions)
        self.table[k] = {"m_tiles": res["m_tiles"], "bin_split": res["bin_split"], "best_ms": res["best_ms"],
                         "default_ms": res["default_ms"]}
        return dict(self.table[k], cached=False)

    def save(self, path: str | None = None):
        path = path or self.path
        if path:
            with open(path, "w") as f:
                json.dump(self.table, f, indent=1, sort_keys=True)
