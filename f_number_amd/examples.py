"""Experimental-data I/O and the examples pipeline of the reference (examples.m).

The reference's examples load ``data_RF.mat`` (examples.m:38) -- variables
``data_RF`` [N_t x N_el x N_angles], ``f_s``, ``f_lb``, ``f_ub``,
``steering_angles_deg``, ``element_width``, ``element_pitch``, ``c_avg`` -- and
beamform it five ways: per steering angle with coherent compounding
(examples.m:56-78), the graphical abstract's three receive F-numbers
(examples.m:104-126), four receive windows (examples.m:144-166), the two
normalizations (examples.m:184-206), and four bandwidths (examples.m:224-249),
each shown as a log-compressed image over a 60 dB dynamic range.

``load_data_rf`` / ``save_data_rf`` read and write that file (MATLAB v4-v7
formats through scipy.io; a v7.3 / HDF5 file must be re-saved with -v7);
``run_examples`` runs the five sections on the GPU beamformer and returns the
images; ``to_db`` is the display transform; ``python -m f_number_amd.examples
data_RF.mat`` runs everything and writes the images (and PNG figures when
matplotlib is present).
"""
from __future__ import annotations

import argparse
import os

import numpy as np

from . import f_numbers, normalizations, windows
from ._lib import FNumberError
from .das import das_pw

VARIABLES = ("data_RF", "f_s", "f_lb", "f_ub", "steering_angles_deg", "element_width", "element_pitch", "c_avg")
SCALARS = ("f_s", "f_lb", "f_ub", "element_width", "element_pitch", "c_avg")


def load_data_rf(path: str) -> dict:
    """The variables of data_RF.mat as numpy values (data_RF keeps its MATLAB layout
    [N_t, N_el, N_angles] and dtype; scalars as float; steering_angles_deg 1-D)."""
    import scipy.io

    try:
        m = scipy.io.loadmat(path, squeeze_me=False)
    except NotImplementedError as e:  # MATLAB v7.3 is HDF5
        raise FNumberError("examples:UnsupportedMatFile",
                           f"{path}: MATLAB v7.3 (HDF5) files are not readable here; re-save with save(..., '-v7')") from e
    missing = [v for v in VARIABLES if v not in m]
    if missing:
        raise FNumberError("examples:MissingVariables", f"{path} lacks {', '.join(missing)}")
    out = {}
    rf = np.asarray(m["data_RF"])
    if rf.ndim == 2:
        rf = rf[:, :, None]
    if rf.ndim != 3 or not np.isrealobj(rf):
        raise FNumberError("examples:InvalidRFData", "data_RF must be real [N_t x N_el x N_angles]")
    out["data_RF"] = rf
    for v in SCALARS:
        a = np.asarray(m[v], dtype=np.float64)
        if a.size != 1:
            raise FNumberError("examples:InvalidScalar", f"{v} must be a scalar")
        out[v] = float(a.reshape(-1)[0])
    out["steering_angles_deg"] = np.asarray(m["steering_angles_deg"], dtype=np.float64).reshape(-1)
    if out["steering_angles_deg"].size != rf.shape[2]:
        raise FNumberError("examples:InvalidSteeringAngles", "one steering angle per RF data set is required")
    return out


def save_data_rf(path: str, data_RF, f_s, f_lb, f_ub, steering_angles_deg, element_width, element_pitch, c_avg):
    """Write a data_RF.mat with the reference's variable names (MATLAB v5 format)."""
    import scipy.io

    scipy.io.savemat(path, {"data_RF": np.asarray(data_RF), "f_s": float(f_s), "f_lb": float(f_lb),
                            "f_ub": float(f_ub),
                            "steering_angles_deg": np.asarray(steering_angles_deg, dtype=np.float64).reshape(1, -1),
                            "element_width": float(element_width), "element_pitch": float(element_pitch),
                            "c_avg": float(c_avg)}, do_compression=False)


def grid(element_pitch: float, n_x: int = 512, n_z: int = 512, z_offset: int = 64):
    """examples.m:49-50: x = (-255.5:255.5) pitch / 4, z = (64 + (0:511)) pitch / 4."""
    x = (np.arange(n_x) - (n_x - 1) / 2.0) * element_pitch / 4
    z = (z_offset + np.arange(n_z)) * element_pitch / 4
    return x, z


def to_db(image, dynamic_range_dB: float = 60.0):
    """20 log10(|image| / max |image|), clipped to [-dynamic_range, 0] (examples.m:86)."""
    a = np.abs(np.asarray(image))
    with np.errstate(divide="ignore"):
        db = 20 * np.log10(a / a.max())
    return np.clip(db, -dynamic_range_dB, 0.0)


def run_examples(data: dict, sections=(1, 2, 3, 4, 5), index_t0: float = 8, n_x: int = 512, n_z: int = 512,
                 index_angle: int = 0, index_gpu: int = 0, semantics: str = "cpu", output_scale: str = "mex") -> dict:
    """The five sections of examples.m on the GPU beamformer.  Returns
    {section: {label: image [N_z, N_x] complex64}} plus 'positions_x' / 'positions_z';
    section 1 also holds 'compound' = the sum of the per-angle images (examples.m:78).
    index_angle selects the angle of sections 2-5 (examples.m: the first one)."""
    from .platforms import gpu

    rf = data["data_RF"]
    ang = np.deg2rad(data["steering_angles_deg"])
    pitch, width, c, f_s = data["element_pitch"], data["element_width"], data["c_avg"], data["f_s"]
    fb = (data["f_lb"], data["f_ub"])
    x, z = grid(pitch, n_x, n_z)
    plat = gpu(index_gpu)
    kw = dict(semantics=semantics, output_scale=output_scale)

    def beam(a, f_bounds=fb, window=None, F=None, norm=None):
        return das_pw(x, z, rf[:, :, a], f_s, float(ang[a]), width, pitch, c, f_bounds, index_t0, window, F, norm,
                      plat, **kw)

    out = {"positions_x": x, "positions_z": z}
    if 1 in sections:  # typical usage: every angle, coherent compounding
        imgs = {f"angle_{d:g}deg": beam(a) for a, d in enumerate(data["steering_angles_deg"])}
        imgs["compound"] = np.sum(np.stack([v for v in imgs.values()]), axis=0)
        out[1] = imgs
    a0 = index_angle
    if 2 in sections:  # graphical abstract: full aperture, fixed, proposed
        out[2] = {"Full aperture": beam(a0, F=f_numbers.constant(0)),
                  "Fixed": beam(a0, F=f_numbers.constant(1.5)),
                  "Proposed": beam(a0, F=f_numbers.grating.angle_lb(60, 1.5))}
    if 3 in sections:  # apodization
        out[3] = {str(w): beam(a0, window=w) for w in (windows.boxcar(), windows.tukey(0.2), windows.triangular(),
                                                       windows.hann())}
    if 4 in sections:  # normalization
        out[4] = {str(n): beam(a0, norm=n) for n in (normalizations.off(), normalizations.on())}
    if 5 in sections:  # bandwidth: f_c (1 +- fb/2)
        f_c = np.mean(fb)
        out[5] = {f"fractional_bandwidth_{int(b * 100)}": beam(a0, f_bounds=tuple(f_c * (1 + np.array([-1, 1]) * b / 2)))
                  for b in (1.0, 0.75, 0.5, 0.25)}
    return out


def save_figures(results: dict, out_dir: str, dynamic_range_dB: float = 60.0) -> list:
    """One PNG per section (log-compressed, gray, examples.m figures 1-6); returns the files."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    x, z = results["positions_x"] * 1e3, results["positions_z"] * 1e3
    files = []
    for sec in (1, 2, 3, 4, 5):
        if sec not in results:
            continue
        imgs = results[sec]
        fig, axs = plt.subplots(1, len(imgs), figsize=(3.2 * len(imgs), 3.6), squeeze=False)
        for ax, (label, img) in zip(axs[0], imgs.items()):
            ax.imshow(to_db(img, dynamic_range_dB), extent=(x[0], x[-1], z[-1], z[0]), cmap="gray",
                      vmin=-dynamic_range_dB, vmax=0, aspect="equal")
            ax.set_title(label, fontsize=8)
            ax.set_xlabel("Lateral position (mm)")
            ax.set_ylabel("Axial position (mm)")
        fig.tight_layout()
        f = os.path.join(out_dir, f"examples_section{sec}.png")
        fig.savefig(f, dpi=110)
        plt.close(fig)
        files.append(f)
    return files


def main(argv=None):
    ap = argparse.ArgumentParser(description="Run the reference's examples (examples.m) on the MI355X beamformer.")
    ap.add_argument("mat", help="data_RF.mat (variables of examples.m:38)")
    ap.add_argument("--out", default="examples_out", help="output directory (images .npz, figures .png)")
    ap.add_argument("--sections", default="1,2,3,4,5")
    ap.add_argument("--gpu", type=int, default=0)
    args = ap.parse_args(argv)
    data = load_data_rf(args.mat)
    res = run_examples(data, sections=tuple(int(s) for s in args.sections.split(",")), index_gpu=args.gpu)
    os.makedirs(args.out, exist_ok=True)
    flat = {f"s{k}_{lab}": v for k, d in res.items() if isinstance(k, int) for lab, v in d.items()}
    np.savez(os.path.join(args.out, "images.npz"), positions_x=res["positions_x"], positions_z=res["positions_z"],
             **flat)
    try:
        files = save_figures(res, args.out)
    except ImportError:
        files = []
    print(f"wrote {len(flat)} images to {args.out}/images.npz" + (f" and {len(files)} figures" if files else ""))


if __name__ == "__main__":
    main()
