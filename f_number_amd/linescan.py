"""Line-by-line scanning with the frequency-dependent F-number.

The reference names line-by-line scanning among its applications (README.md:31-33,
its reference [3]) but ships no code for it.  Here line i of the image, at lateral
position x_i, is formed from transmit event i alone:

* transmit: a wave focused at (x_f,i, z_f) -- a virtual source below the array,
  with the transmit time of flight (|r_vs| + sign(z - z_vs) |r - r_vs|) / c of
  fnumber_hip.h fn_pw_params.virtual_sources;
* receive: the Fourier-domain beamformer with its frequency-dependent F-number
  (das_pw.m:302-485 with that transmit time of flight) on the grid column (x_i, z).

All lines are one plan in line-scan mode (fn_pw_params.line_scan): the spectra of
every transmit event in one launch, and one beamforming launch in which the
workgroups of column i contract transmit row i only.  `LineScan` holds that plan
for device buffers, `das_lbl` is the host-array form (MATLAB layout).
"""
from __future__ import annotations

import numpy as np

from . import f_numbers, normalizations, platforms, windows
from ._lib import FNumberError
from .plan import PwPlan

PREFIX = "das_lbl"


def _err(ident: str, msg: str):
    raise FNumberError(f"{PREFIX}:{ident}", msg)


class LineScan:
    """Plan of a line-by-line scan: line i at positions_x[i] from transmit i focused at
    (focus_x[i], focus_z) (focus_x defaults to positions_x).  RF (N_lines, N_el, N_t)
    float32/float64 on the device, image (N_lines, N_z) complex64 on the device."""

    def __init__(self, positions_x, positions_z, n_samples, n_elements, f_s, element_width, element_pitch, c_0,
                 f_bounds, focus_z, focus_x=None, index_t0=0.0, window=None, F_number=None, normalization=None,
                 index_gpu=0, semantics="cpu", output_scale="mex"):
        px = np.ascontiguousarray(np.asarray(positions_x, dtype=np.float64).reshape(-1))
        pz = np.ascontiguousarray(np.asarray(positions_z, dtype=np.float64).reshape(-1))
        if px.size < 1 or pz.size < 1:
            _err("InvalidPositions", "positions_x and positions_z must hold at least one value each")
        fx = px if focus_x is None else np.asarray(focus_x, dtype=np.float64).reshape(-1)
        if fx.size != px.size:
            _err("InvalidFocus", "one focus_x per line (per entry of positions_x)")
        fz = float(focus_z)
        if not (np.isfinite(fz) and fz > 0.0) or not np.all(np.isfinite(fx)):
            _err("InvalidFocus", "focus_z must be a finite depth below the array (> 0)")
        window = window if window is not None else windows.tukey(0.2)
        F_number = F_number if F_number is not None else f_numbers.grating.angle_lb(60, 1.5)
        normalization = normalization if normalization is not None else normalizations.on()
        self.positions_x, self.positions_z, self.focus_x, self.focus_z = px, pz, fx, fz
        self.n_samples, self.n_elements = int(n_samples), int(n_elements)
        self.plan = PwPlan(px, pz, n_samples, n_elements, np.zeros(px.size), f_s, element_width, element_pitch, c_0,
                           f_bounds, index_t0, window, F_number, normalization, index_gpu, semantics, "fp32",
                           output_scale, virtual_sources=np.stack([fx, np.full(px.size, fz)], axis=1),
                           line_scan=True)

    @property
    def n_lines(self) -> int:
        return self.positions_x.size

    def execute(self, rf, image, weights=None, stream=None):
        """rf: (N_lines, N_el, N_t) device tensor; image: (N_lines, N_z) complex64 device tensor;
        weights: optional float32 (N_lines, N_z), the receive window sum of the last bin."""
        if tuple(rf.shape) != (self.n_lines, self.n_elements, self.n_samples):
            _err("InvalidRFData", f"RF must be a ({self.n_lines}, {self.n_elements}, {self.n_samples}) tensor")
        if tuple(image.shape) != (self.n_lines, self.positions_z.size):
            _err("InvalidArgument", f"image must be a ({self.n_lines}, {self.positions_z.size}) tensor")
        self.plan.execute(rf, image, weights=weights, stream=stream)

    def autotune(self, rf, image, repetitions: int = 3, stream=None) -> dict:
        """Pick the bin split of the beamforming launch by timing (image is overwritten)."""
        return self.plan.autotune(rf, image, repetitions, stream)

    def info(self) -> dict:
        return self.plan.info()

    def close(self):
        self.plan.close()


def das_lbl(positions_x, positions_z, data_RF, f_s, focus_z, element_width, element_pitch, c_0, f_bounds=None,
            index_t0=0.0, window=None, F_number=None, normalization=None, platform=None, *, focus_x=None,
            semantics: str = "cpu", output_scale: str = "mex"):
    """Line-by-line scan from host data: data_RF [N_t, N_el, N_lines] (MATLAB layout, one
    transmit event per line, line i focused at (focus_x[i], focus_z)) -> image [N_z, N_lines]
    complex64 in the das_pw output scale."""
    import torch

    rf = np.asarray(data_RF)
    if rf.ndim == 2:
        rf = rf[:, :, None]
    if rf.ndim != 3 or not np.isrealobj(rf) or rf.dtype not in (np.float32, np.float64):
        _err("InvalidRFData", "data_RF must be real float32/float64 [N_t x N_el x N_lines]")
    n_t, n_el, n_lines = rf.shape
    if np.asarray(positions_x).size != n_lines:
        _err("InvalidRFData", "one transmit event (third dimension of data_RF) per entry of positions_x")
    if f_bounds is None or np.size(f_bounds) == 0:
        f_bounds = (np.finfo(float).eps, float(f_s) / 2)  # the das_pw default
    plat = platform if platform is not None else platforms.gpu(0)
    if not isinstance(plat, platforms.platform):
        _err("NoPlatform", "platform must be scalar platforms.platform!")
    if not isinstance(plat, platforms.gpu):
        _err("UnsupportedPlatform", "only platforms.gpu is served")
    index_gpu = plat.index
    scan = LineScan(positions_x, positions_z, n_t, n_el, f_s, element_width, element_pitch, c_0, f_bounds, focus_z,
                    focus_x, index_t0, window, F_number, normalization, index_gpu, semantics, output_scale)
    try:
        dev = torch.device(f"cuda:{index_gpu}")
        rf_d = torch.from_numpy(np.ascontiguousarray(np.transpose(rf, (2, 1, 0)))).to(dev)
        img = torch.empty((n_lines, np.asarray(positions_z).size), dtype=torch.complex64, device=dev)
        scan.execute(rf_d, img)
        torch.cuda.synchronize(dev)
        return img.cpu().numpy().T
    finally:
        scan.close()
