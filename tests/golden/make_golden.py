"""Write the golden fixtures of this directory from the FP64 oracle (oracle/).

The reference ships no test vectors (no tests; its data_RF.mat is missing), so
these fixtures are the oracle's own outputs on small seeded inputs: F-number
tables, window samples, small plane-wave / synthetic-aperture images and the
differentiation quantities.  tests/test_golden.py checks that the oracle still
reproduces them (CPU) and that the HIP path matches them (GPU).  Inputs are
stored in the fixtures themselves (numpy RNG, not the product's synthesiser).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

# array and acquisition: 128-element linear array, 0.3048 mm pitch, 0.27 mm width, 20.832 MHz, 1540 m/s
GEO = dict(f_s=20.832e6, width=0.27e-3, pitch=0.3048e-3, c_0=1540.0, f_bounds=(2.25e6, 6.75e6))

F_KINDS = {
    "constant_1.5": ("constant", 1.5),
    "angle_lb_60_1.5_10": ("angle_lb", 60, 1.5, 10),
    "angle_lb_70_1.5_10": ("angle_lb", 70, 1.5, 10),
    "distance_constant_10_2": ("distance_constant", 10, 2.0),
    "szabo_0.918": ("szabo", 0.918),
    "directivity_soft_0.918_6": ("directivity_soft", 0.918, 6),
    "directivity_rigid_0.918_6": ("directivity_rigid", 0.918, 6),
    "perrot_0.918": ("perrot", 0.918),
    "distance_none": ("distance_none",),
}

# plane-wave image cases: (window, parameter), F-number, normalization, semantics
PW_CASES = {
    "tukey_anglelb_on_cpu": (("tukey", 0.5), ("angle_lb", 60, 1.5, 10), "on", "cpu"),
    "boxcar_const_off_mex": (("boxcar", 0.0), ("constant", 1.5), "off", "mex"),
    "hann_anglelb_on_cpu": (("hann", 0.0), ("angle_lb", 60, 1.5, 10), "on", "cpu"),
    "triangular_anglelb_on_mex": (("triangular", 0.0), ("angle_lb", 60, 1.5, 10), "on", "mex"),
}


def pw_inputs():
    rng = np.random.default_rng(20261202)
    x = np.linspace(-2e-3, 2e-3, 8)
    z = np.linspace(3e-3, 8e-3, 6)
    angles = np.deg2rad(np.array([-4.0, 0.0, 4.0]))
    rf = rng.normal(size=(3, 16, 256)).astype(np.float32)  # [n_tx, n_el, n_t]
    return x, z, angles, rf


def sa_inputs():
    rng = np.random.default_rng(20261203)
    x = np.linspace(-1e-3, 1e-3, 6)
    z = np.linspace(3e-3, 7e-3, 5)
    rf = rng.normal(size=(8, 8, 256)).astype(np.float32)  # [n_tx, n_rx, n_t]
    return x, z, rf


def compute():
    out = {}
    # F-number tables over the element pitch in wavelengths of the cfg0 band (N = 2509, bins 271..812)
    N, lb, ub = O.grid(2048, 128, GEO["pitch"], 5e-3, GEO["f_s"], GEO["c_0"], *GEO["f_bounds"])
    r = GEO["pitch"] * np.arange(lb, ub + 1) * GEO["f_s"] / N / GEO["c_0"]
    out["fnumbers"] = {"r": r, **{k: O.fnumber_values(O.fnumber(*v), r) for k, v in F_KINDS.items()}}
    # window samples at t = |x| / half-width
    t = np.linspace(0.0, 1.5, 61)
    win = {"t": t}
    for sem in ("cpu", "mex"):
        for name, p in (("boxcar", 0.0), ("hann", 0.0), ("tukey", 0.5), ("tukey", 0.2), ("triangular", 0.0)):
            win[f"{name}_{p}_{sem}"] = O.window(name, p, t, sem)
    out["windows"] = win
    # plane-wave images (CPU output scale)
    x, z, angles, rf = pw_inputs()
    pw = {"x": x, "z": z, "angles": angles, "rf": rf}
    for key, (w, F, norm, sem) in PW_CASES.items():
        img, _ = O.das_pw(x, z, rf, GEO["f_s"], angles, GEO["width"], GEO["pitch"], GEO["c_0"], GEO["f_bounds"], 0.0,
                          w, O.fnumber(*F), norm, sem)
        pw[key] = img
    out["das_pw"] = pw
    # synthetic aperture image
    xs, zs, rfs = sa_inputs()
    img, _ = O.das_sa(xs, zs, rfs, GEO["f_s"], GEO["width"], GEO["pitch"], GEO["c_0"], GEO["f_bounds"], 0.0,
                      window_tx=("tukey", 0.5), window_rx=("tukey", 0.5))
    out["das_sa"] = {"x": xs, "z": zs, "rf": rfs, "tukey_anglelb_on_cpu": img}
    # differentiation quantities of the first plane-wave case at c_tof = 1510 m/s with delay offsets
    rng = np.random.default_rng(20261204)
    delays = rng.normal(0.0, 10e-9, 16)
    y = rng.normal(size=(z.size, x.size)) + 1j * rng.normal(size=(z.size, x.size))
    d = O.das_pw_diff(x, z, rf, GEO["f_s"], angles, GEO["width"], GEO["pitch"], GEO["c_0"], GEO["f_bounds"], 0.0,
                      ("tukey", 0.5), O.fnumber("angle_lb", 60, 1.5, 10), "on", "cpu", c_tof=1510.0,
                      delay_offsets=delays, g=y, y=y, want_dc=True)
    out["diff"] = {"c_tof": np.array(1510.0), "delays": delays, "y": y, "image": d["image"], "dIdc": d["dIdc"],
                   "vjp_delay": d["vjp_delay"], "vjp_c": np.array(d["vjp_c"]), "adjoint": d["adjoint"]}
    return out


def main():
    O.build()
    for name, arrays in compute().items():
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        print("wrote", name)


if __name__ == "__main__":
    main()
