"""Seeded synthetic acquisitions and the workload presets they are used with.

The reference ships no data (data_RF.mat is missing), so every workload is
generated: a linear array of 128 elements (pitch 0.3048 mm, width 0.27 mm),
c = 1540 m/s, f_s = 20.832 MHz, point plus speckle scatterers with a 5 MHz
Gaussian-modulated pulse of 60 % bandwidth and 30 dB SNR (libfnumber_hip's
fn_synth_rf).  Seeds: 20260000 + workload index (noise: + 1000).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib

PITCH = 0.3048e-3
WIDTH = 0.27e-3
C0 = 1540.0
FS = 20.832e6
N_EL = 128


@dataclass
class Workload:
    name: str
    mode: str                      # "pw" or "sa"
    n_samples: int
    angles_deg: np.ndarray         # PW steering angles (deg)
    positions_x: np.ndarray
    positions_z: np.ndarray
    f_bounds: tuple
    n_points: int = 20
    n_speckle: int = 20000
    seed: int = 20260000
    n_elements: int = N_EL
    pitch: float = PITCH
    width: float = WIDTH
    c_0: float = C0
    f_s: float = FS
    scatter_x: tuple = (-19e-3, 19e-3)
    scatter_z: tuple = (5e-3, 40e-3)
    snr_db: float = 30.0
    extra: dict = field(default_factory=dict)
    # medium: (thickness m, speed of sound m/s) of a surface layer; (0, 0) = homogeneous
    layer: tuple = (0.0, 0.0)
    # per-element receive delay offsets (s) of the acquisition, or None
    delay_offsets: np.ndarray | None = None
    # flow tube (x, z, radius, angle rad, v_max m/s) and the frame time (s)
    tube: tuple = (0.0, 0.0, 0.0, 0.0, 0.0)
    frame_time: float = 0.0
    # virtual-source transmits (x_vs, z_vs) in m, one per entry of angles_deg (PW mode), or None
    virtual_sources: np.ndarray | None = None

    @property
    def angles(self) -> np.ndarray:
        return np.deg2rad(self.angles_deg)

    @property
    def n_tx(self) -> int:
        return self.angles_deg.size if self.mode == "pw" else self.n_elements

    @property
    def n_voxels(self) -> int:
        return self.positions_x.size * self.positions_z.size

    def synth_params(self, on_device: bool = False, index_gpu: int = 0):
        P = _lib.SynthParams()
        P.struct_size = C.sizeof(_lib.SynthParams)
        P.seed = self.seed
        P.n_points, P.n_speckle = self.n_points, self.n_speckle
        P.point_amplitude = 10.0
        P.x_min, P.x_max = self.scatter_x
        P.z_min, P.z_max = self.scatter_z
        P.n_samples, P.n_elements = self.n_samples, self.n_elements
        P.f_s, P.element_pitch, P.c_0 = self.f_s, self.pitch, self.c_0
        P.f_c, P.fractional_bandwidth = 5e6, 0.6
        P.snr_db = self.snr_db
        P.mode = (2 if self.virtual_sources is not None else 0) if self.mode == "pw" else 1
        ang = np.ascontiguousarray(self.angles, dtype=np.float64)
        P.n_angles = ang.size if self.mode == "pw" else 0
        P.steering_angles = ang.ctypes.data_as(C.POINTER(C.c_double))
        P.on_device = int(on_device)
        P.index_gpu = index_gpu
        P.layer_thickness, P.layer_c = float(self.layer[0]), float(self.layer[1])
        keep = [ang]
        if self.delay_offsets is not None:
            d = np.ascontiguousarray(self.delay_offsets, dtype=np.float64)
            assert d.size == self.n_elements
            P.delay_offsets = d.ctypes.data_as(C.POINTER(C.c_double))
            keep.append(d)
        P.tube_x, P.tube_z, P.tube_radius, P.tube_angle, P.tube_v_max = (float(v) for v in self.tube)
        P.frame_time = float(self.frame_time)
        if self.virtual_sources is not None:
            vs = np.ascontiguousarray(np.asarray(self.virtual_sources, dtype=np.float64).reshape(-1, 2))
            assert vs.shape[0] == ang.size
            P.virtual_sources = vs.ctypes.data_as(C.POINTER(C.c_double))
            keep.append(vs)
        return P, keep

    def rf_host(self) -> np.ndarray:
        """RF as float32 [N_tx, N_el, N_t] (C order) = MATLAB's [N_t x N_el x N_tx]."""
        P, keep = self.synth_params(False)
        out = np.empty((self.n_tx, self.n_elements, self.n_samples), dtype=np.float32)
        _lib.check(_lib.load().fn_synth_rf(C.byref(P), out.ctypes.data), "fn_synth_rf")
        return out

    def rf_device(self, device="cuda:0"):
        """The same acquisition generated on the GPU into a torch tensor (N_tx, N_el, N_t)."""
        import torch

        dev = torch.device(device)
        out = torch.empty((self.n_tx, self.n_elements, self.n_samples), dtype=torch.float32, device=dev)
        P, keep = self.synth_params(True, dev.index or 0)
        _lib.check(_lib.load().fn_synth_rf(C.byref(P), C.c_void_p(out.data_ptr())), "fn_synth_rf")
        return out

    def scatterers(self):
        P, keep = self.synth_params(False)
        n = self.n_points + self.n_speckle
        xs, zs, amp = np.empty(n), np.empty(n), np.empty(n)
        dp = C.POINTER(C.c_double)
        _lib.check(_lib.load().fn_synth_scatterers(C.byref(P), xs.ctypes.data_as(dp), zs.ctypes.data_as(dp),
                                                   amp.ctypes.data_as(dp)), "fn_synth_scatterers")
        return xs, zs, amp


def _lin(lo, hi, n):
    return np.linspace(lo, hi, n)


def cpwc_small() -> Workload:
    """Plain CPWC, 11 angles -10..10 deg, 2048 samples, 2.25-6.75 MHz, 256 x 256 voxels over
    x in [-19, 19] mm, z in [5, 40] mm (proposed F-number, Tukey(0.5), normalization on)."""
    return Workload("cpwc_11ang_256x256", "pw", 2048, np.arange(-10.0, 10.0 + 1e-9, 2.0),
                    _lin(-19e-3, 19e-3, 256), _lin(5e-3, 40e-3, 256), (2.25e6, 6.75e6), seed=20260000)


def cpwc_full() -> Workload:
    """Plain CPWC at full size: 75 angles -18.5..18.5 deg (0.5 deg), 3328 samples, 1-9 MHz,
    1024 axial (0.035 mm from 5 mm) x 512 lateral (0.075 mm, centred) voxels, 50 points."""
    x = (np.arange(512) - 255.5) * 0.075e-3
    z = 5e-3 + np.arange(1024) * 0.035e-3
    return Workload("cpwc_75ang_1024x512", "pw", 3328, np.arange(-18.5, 18.5 + 1e-9, 0.5), x, z,
                    (1e6, 9e6), n_points=50, seed=20260001)


def saft_full() -> Workload:
    """Full synthetic aperture: 128 x 128 x 3000 samples, 512 x 512 voxels at 0.075 mm,
    2.25-6.75 MHz."""
    x = (np.arange(512) - 255.5) * 0.075e-3
    z = 5e-3 + np.arange(512) * 0.075e-3
    return Workload("saft_128x128_512x512", "sa", 3000, np.zeros(0), x, z, (2.25e6, 6.75e6), seed=20260002)


def diverging_waves(n_sources: int = 11, z_vs: float = -15e-3) -> Workload:
    """Diverging-wave compounding: n_sources virtual sources at z_vs behind the array,
    x_vs spread over [-10, 10] mm, 2048 samples, 2.25-6.75 MHz, 256 x 256 voxels over
    x in [-19, 19] mm, z in [5, 40] mm."""
    w = Workload("diverging_11vs_256x256", "pw", 2048, np.zeros(n_sources), _lin(-19e-3, 19e-3, 256),
                 _lin(5e-3, 40e-3, 256), (2.25e6, 6.75e6), seed=20260005)
    w.virtual_sources = np.stack([np.linspace(-10e-3, 10e-3, n_sources), np.full(n_sources, z_vs)], axis=1)
    return w


def line_scan(n_lines: int = 128, n_z: int = 256, focus_z: float = 20e-3) -> Workload:
    """Line-by-line scan: transmit i focused at (x_i, focus_z), x_i = positions_x[i] spread
    over [-19, 19] mm; 2048 samples, 2.25-6.75 MHz, n_z depths over z in [5, 40] mm.  RF
    (N_lines, N_el, N_t) holds one transmit event per image line."""
    w = Workload(f"line_scan_{n_lines}x{n_z}", "pw", 2048, np.zeros(n_lines), _lin(-19e-3, 19e-3, n_lines),
                 _lin(5e-3, 40e-3, n_z), (2.25e6, 6.75e6), seed=20260006)
    w.virtual_sources = np.stack([w.positions_x, np.full(n_lines, focus_z)], axis=1)
    return w


def autofocus_layer() -> Workload:
    """Differentiation-form acquisition: the 11-angle array of cpwc_small (2048 samples,
    2.25-6.75 MHz) over a 6 mm surface layer at 1450 m/s on a 1540 m/s background, with
    per-element receive delay offsets N(0, 20 ns) (seed 20260004 + 2000); 384 x 384 voxels
    over x in [-19, 19] mm, z in [5, 40] mm."""
    w = Workload("autofocus_layer_11ang_384x384", "pw", 2048, np.arange(-10.0, 10.0 + 1e-9, 2.0),
                 _lin(-19e-3, 19e-3, 384), _lin(5e-3, 40e-3, 384), (2.25e6, 6.75e6), seed=20260004)
    w.layer = (6e-3, 1450.0)
    w.delay_offsets = np.random.default_rng(w.seed + 2000).normal(0.0, 20e-9, w.n_elements)
    return w


def doppler_frame(frame: int, frame_interval: float = 2e-4) -> Workload:
    """Ultrafast-Doppler ensemble frame: 5 angles -4..4 deg, 1536 samples, 2.25-6.75 MHz,
    256 x 256 voxels over x in [-19, 19] mm, z in [5, 40] mm; scatterers inside a 4 mm tube
    (centre z = 20 mm, axis tilted 10 deg) flow at up to 0.5 m/s (parabolic profile);
    frame k is the medium at time k * frame_interval (seed 20260003)."""
    w = Workload("doppler_5ang_256x256", "pw", 1536, np.arange(-4.0, 4.0 + 1e-9, 2.0),
                 _lin(-19e-3, 19e-3, 256), _lin(5e-3, 40e-3, 256), (2.25e6, 6.75e6), seed=20260003)
    w.tube = (0.0, 20e-3, 2e-3, np.deg2rad(10.0), 0.5)
    w.frame_time = frame * frame_interval
    return w


def scaled(w: Workload, n_x: int, n_z: int, n_angles: int | None = None, n_samples: int | None = None,
           n_speckle: int | None = None, n_elements: int | None = None) -> Workload:
    """A smaller copy of a workload (same extent, fewer voxels/angles) for parity tests."""
    import copy

    v = copy.deepcopy(w)
    v.name = f"{w.name}_scaled"
    v.positions_x = np.linspace(w.positions_x[0], w.positions_x[-1], n_x)
    v.positions_z = np.linspace(w.positions_z[0], w.positions_z[-1], n_z)
    if n_angles is not None and w.mode == "pw":
        idx = np.round(np.linspace(0, w.angles_deg.size - 1, n_angles)).astype(int)
        v.angles_deg = w.angles_deg[idx]
        if w.virtual_sources is not None:
            v.virtual_sources = np.asarray(w.virtual_sources)[idx]
    if n_samples is not None:
        v.n_samples = n_samples
    if n_speckle is not None:
        v.n_speckle = n_speckle
    if n_elements is not None:
        v.n_elements = n_elements
    return v
