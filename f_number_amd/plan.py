"""Device-resident execution: plans over libfnumber_hip's fn_plan_* entry points.

A plan validates the configuration once, keeps the grid, F-number tables and
the spectra workspace on the GPU, and executes on device pointers.  Device
buffers are torch tensors (torch is used as the allocator/stream provider):
RF as a contiguous tensor shaped (N_tx, N_el, N_t) -- the memory of MATLAB's
column-major [N_t x N_el x N_tx] -- and the image as complex64 (N_x, N_z).
"""
from __future__ import annotations

import ctypes as C

from . import _lib, f_numbers, normalizations, windows
from .das import pw_params, sa_params


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class _Plan:
    prefix = "fn_plan"

    def __init__(self):
        self._h = C.c_void_p()
        self._keep = None

    def info(self) -> dict:
        inf = _lib.PlanInfo()
        _lib.check(_lib.load().fn_plan_info_get(self._h, C.byref(inf)), self.prefix)
        return {f: getattr(inf, f) for f, _ in _lib.PlanInfo._fields_}

    def execute(self, rf, image, weights=None, weights_rx=None, stream=None):
        """Enqueue the beamformer on `stream` (torch stream or raw handle; None = the
        current torch stream)."""
        import torch

        if not (rf.is_cuda and image.is_cuda):
            raise _lib.FNumberError(f"{self.prefix}:InvalidArgument", "rf and image must be device tensors")
        if not (rf.is_contiguous() and image.is_contiguous()):
            raise _lib.FNumberError(f"{self.prefix}:InvalidArgument", "rf and image must be contiguous")
        if rf.dtype == torch.float32:
            dt = _lib.DTYPE["float32"]
        elif rf.dtype == torch.float64:
            dt = _lib.DTYPE["float64"]
        else:
            raise _lib.FNumberError(f"{self.prefix}:InvalidRFData", "RF must be float32 or float64")
        if image.dtype != torch.complex64:
            raise _lib.FNumberError(f"{self.prefix}:InvalidArgument", "image must be complex64")
        inf = self.info()
        if image.numel() != inf["n_voxels"] * getattr(self, "n_frames", 1):
            raise _lib.FNumberError(f"{self.prefix}:InvalidArgument", "image size does not match the grid")
        if rf.numel() != self._rf_numel:
            raise _lib.FNumberError(f"{self.prefix}:InvalidRFData", "RF size does not match the plan")
        for w in (weights, weights_rx):
            if w is not None and (w.dtype != torch.float32 or w.numel() != inf["n_voxels"] or not w.is_cuda):
                raise _lib.FNumberError(f"{self.prefix}:InvalidArgument", "weights must be float32 device tensors")
        if stream is None:
            stream = torch.cuda.current_stream(rf.device)
        handle = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
        _lib.check(_lib.load().fn_plan_execute(self._h, _ptr(rf), dt, _ptr(image), _ptr(weights), _ptr(weights_rx),
                                               C.c_void_p(handle)), self.prefix)

    def autotune(self, rf, image, repetitions: int = 3, stream=None) -> dict:
        """Time the candidate tilings (transmit tiles per workgroup, bin split) on rf/image
        (image contents are overwritten) and keep the fastest; returns the choice."""
        r = _lib.TuneResult()
        _lib.check(_lib.load().fn_plan_autotune(self._h, _ptr(rf), _rf_dtype(rf, self.prefix), _ptr(image),
                                                int(repetitions), C.byref(r), _stream(stream, rf.device)),
                   self.prefix)
        return {f: getattr(r, f) for f, _ in _lib.TuneResult._fields_}

    def focused_signal(self, stream=None, device=None):
        """Focused RF signal of the last voxel (das_pw.m:477-481 / das_sa.m:501-505) from the
        spectra of the last execute: torch complex64 [N_dft] (CPU path scale)."""
        import torch

        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        out = torch.empty(self.info()["N_dft"], dtype=torch.complex64, device=dev)
        _lib.check(_lib.load().fn_plan_focused_signal(self._h, _ptr(out), _stream(stream, dev)), self.prefix)
        return out

    def set_tiling(self, m_tiles: int, bin_split: int):
        _lib.check(_lib.load().fn_plan_set_tiling(self._h, int(m_tiles), int(bin_split)), self.prefix)

    def set_profiling(self, enable: bool = True):
        _lib.check(_lib.load().fn_plan_set_profiling(self._h, int(bool(enable))), self.prefix)

    def get_profile(self):
        """(spectra_ms_total, beamform_ms_total, runs) since the last call (HIP events)."""
        a, b, n = C.c_double(), C.c_double(), C.c_int32()
        _lib.check(_lib.load().fn_plan_get_profile(self._h, C.byref(a), C.byref(b), C.byref(n)), self.prefix)
        return a.value, b.value, n.value

    def close(self):
        if self._h:
            _lib.load().fn_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _rf_dtype(rf, prefix):
    import torch

    if rf.dtype == torch.float32:
        return _lib.DTYPE["float32"]
    if rf.dtype == torch.float64:
        return _lib.DTYPE["float64"]
    raise _lib.FNumberError(f"{prefix}:InvalidRFData", "RF must be float32 or float64")


def _stream(stream, device):
    import torch

    if stream is None:
        stream = torch.cuda.current_stream(device)
    return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


def _need(t, dtype, numel, what, prefix):
    if t is None or not t.is_cuda or not t.is_contiguous() or t.dtype != dtype or t.numel() != numel:
        raise _lib.FNumberError(f"{prefix}:InvalidArgument",
                                f"{what} must be a contiguous {dtype} device tensor of {numel} elements")


class PwPlan(_Plan):
    """Plan of das_pw for a fixed grid, acquisition and angle set.

    Differentiation form: set_tof, execute_tangent, adjoint and vjp (fnumber_hip.h)."""
    prefix = "gpu_bf_das_pw_rf"

    def __init__(self, positions_x, positions_z, n_samples, n_elements, steering_angles, f_s, element_width,
                 element_pitch, c_0, f_bounds, index_t0=0.0, window=None, F_number=None, normalization=None,
                 index_gpu=0, semantics="cpu", precision="fp32", output_scale="mex", n_frames=1,
                 virtual_sources=None, line_scan=False):
        """n_frames > 1: an ensemble of frames with this geometry -- RF (n_frames, N_a, N_el, N_t),
        images (n_frames, N_x, N_z), each frame compounding its own angles.
        virtual_sources: (N_a, 2) (x_vs, z_vs) in m per transmit instead of plane waves
        (z_vs < 0 diverging, z_vs > 0 focused; steering_angles then only gives N_a).
        line_scan: column i of the grid is formed from transmit i alone (N_a == N_x;
        fnumber_hip.h fn_pw_params.line_scan), FP32 forward only."""
        super().__init__()
        window = window if window is not None else windows.tukey(0.2)
        F_number = F_number if F_number is not None else f_numbers.grating.angle_lb(60, 1.5)
        normalization = normalization if normalization is not None else normalizations.on()
        P, keep, info = pw_params(positions_x, positions_z, n_samples, n_elements, steering_angles, f_s,
                                  element_width, element_pitch, c_0, f_bounds, index_t0, window, F_number,
                                  normalization, index_gpu, semantics, precision, output_scale, n_frames=n_frames,
                                  virtual_sources=virtual_sources, line_scan=line_scan)
        _lib.check(_lib.load().fn_plan_pw_create(C.byref(P), C.byref(self._h)), self.prefix)
        self._keep = keep
        self.grid = info
        self.n_samples, self.n_elements, self.n_angles = int(n_samples), int(n_elements), len(keep[2])
        self.n_frames = max(1, int(n_frames))
        self.c_0 = float(c_0)
        self.precision = precision
        self.window_class = type(window).__name__
        self._rf_numel = self.n_samples * self.n_elements * self.n_angles * self.n_frames

    def set_tof(self, c_tof=0.0, delay_offsets=None):
        """Speed of sound of the time of flight (0: the plan's c_0) and per-element
        receive delay offsets in seconds (None: zero); apodization stays at c_0."""
        import numpy as np

        d = None
        if delay_offsets is not None:
            d = np.ascontiguousarray(delay_offsets, dtype=np.float64)
            if d.size != self.n_elements:
                raise _lib.FNumberError(f"{self.prefix}:InvalidArgument", "one delay offset per element")
        ptr = d.ctypes.data_as(C.POINTER(C.c_double)) if d is not None else None
        _lib.check(_lib.load().fn_plan_set_tof(self._h, float(c_tof), ptr), self.prefix)

    def execute_tangent(self, rf, image, dimage, weights=None, stream=None):
        """image and dI/dc_tof (complex64 (N_x, N_z) each)."""
        import torch

        n = self.info()["n_voxels"]
        _need(image, torch.complex64, n, "image", self.prefix)
        _need(dimage, torch.complex64, n, "dimage", self.prefix)
        if rf.numel() != self._rf_numel or not rf.is_cuda or not rf.is_contiguous():
            raise _lib.FNumberError(f"{self.prefix}:InvalidRFData", "RF size does not match the plan")
        if weights is not None:
            _need(weights, torch.float32, n, "weights", self.prefix)
        _lib.check(_lib.load().fn_plan_execute_tangent(self._h, _ptr(rf), _rf_dtype(rf, self.prefix), _ptr(image),
                                                       _ptr(dimage), _ptr(weights), _stream(stream, rf.device)),
                   self.prefix)

    def adjoint(self, y, rf_adj=None, stream=None):
        """A^H y: y complex64 (N_x, N_z) -> float32 (N_a, N_el, N_t)."""
        import torch

        _need(y, torch.complex64, self.info()["n_voxels"], "y", self.prefix)
        if rf_adj is None:
            rf_adj = torch.empty((self.n_angles, self.n_elements, self.n_samples), dtype=torch.float32,
                                 device=y.device)
        _need(rf_adj, torch.float32, self._rf_numel, "rf_adj", self.prefix)
        _lib.check(_lib.load().fn_plan_adjoint(self._h, _ptr(y), _ptr(rf_adj), _stream(stream, y.device)),
                   self.prefix)
        return rf_adj

    def vjp(self, rf, g, dimage=None, stream=None):
        """(dJ/d delay [N_el] float64, dJ/dc_tof [1] float64 or None) for cotangent g
        (dJ = sum Re(conj(g) dI)); dJ/dc needs dimage from execute_tangent."""
        import torch

        n = self.info()["n_voxels"]
        _need(g, torch.complex64, n, "g", self.prefix)
        if dimage is not None:
            _need(dimage, torch.complex64, n, "dimage", self.prefix)
        if rf.numel() != self._rf_numel or not rf.is_cuda or not rf.is_contiguous():
            raise _lib.FNumberError(f"{self.prefix}:InvalidRFData", "RF size does not match the plan")
        vd = torch.empty(self.n_elements, dtype=torch.float64, device=g.device)
        vc = torch.empty(1, dtype=torch.float64, device=g.device) if dimage is not None else None
        _lib.check(_lib.load().fn_plan_vjp(self._h, _ptr(rf), _rf_dtype(rf, self.prefix), _ptr(g), _ptr(dimage),
                                           _ptr(vd), _ptr(vc), _stream(stream, g.device)), self.prefix)
        return vd, vc


class SaPlan(_Plan):
    """Plan of das_sa for a fixed grid and acquisition."""
    prefix = "gpu_bf_saft_rf"

    def __init__(self, positions_x, positions_z, n_samples, n_elements, f_s, element_width, element_pitch, c_0,
                 f_bounds, index_t0=0.0, window_tx=None, window_rx=None, F_number_tx=None, F_number_rx=None,
                 normalization_tx=None, normalization_rx=None, index_gpu=0, semantics="cpu", precision="fp32",
                 output_scale="mex"):
        super().__init__()
        window_tx = window_tx if window_tx is not None else windows.tukey(0.2)
        window_rx = window_rx if window_rx is not None else window_tx
        F_number_tx = F_number_tx if F_number_tx is not None else f_numbers.grating.angle_lb(60, 1.5)
        F_number_rx = F_number_rx if F_number_rx is not None else F_number_tx
        normalization_tx = normalization_tx if normalization_tx is not None else normalizations.on()
        normalization_rx = normalization_rx if normalization_rx is not None else normalization_tx
        P, keep, info = sa_params(positions_x, positions_z, n_samples, n_elements, f_s, element_width,
                                  element_pitch, c_0, f_bounds, index_t0, window_tx, window_rx, F_number_tx,
                                  F_number_rx, normalization_tx, normalization_rx, index_gpu, semantics,
                                  precision, output_scale)
        _lib.check(_lib.load().fn_plan_sa_create(C.byref(P), C.byref(self._h)), self.prefix)
        self._keep = keep
        self.grid = info
        self._rf_numel = int(n_samples) * int(n_elements) * int(n_elements)


class Batch:
    """Many plans executed as one hipGraph (fnumber_hip.h fn_batch_*): the launch
    sequences of all plans are captured once over `n_streams` forked streams and
    replayed by one graph launch per execute(); a change of any buffer pointer
    captures again."""

    def __init__(self, plans, n_streams: int = 4, persistent: bool = False, stacking: bool = True):
        self.plans = list(plans)
        arr = (C.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        self._h = C.c_void_p()
        _lib.check(_lib.load().fn_batch_create(arr, len(self.plans), int(n_streams), C.byref(self._h)), "fn_batch")
        self.set_persistent(persistent)
        self.set_stacking(stacking)

    def set_stacking(self, enable: bool = True):
        """Plane-wave plans that differ only in steering angles and RF beamformed by one
        launch over their stacked transmit rows (fnumber_hip.h fn_batch_set_stacking)."""
        _lib.check(_lib.load().fn_batch_set_stacking(self._h, int(bool(enable))), "fn_batch")

    def set_persistent(self, enable: bool = True):
        """Plans of one kernel class as one persistent launch over all their work items
        (fnumber_hip.h fn_batch_set_persistent); False: one launch per plan."""
        _lib.check(_lib.load().fn_batch_set_persistent(self._h, int(bool(enable))), "fn_batch")

    def execute(self, rfs, images, stream=None):
        n = len(self.plans)
        if len(rfs) != n or len(images) != n:
            raise _lib.FNumberError("fn_batch:InvalidArgument", "one RF and one image per plan")
        for p, rf, img in zip(self.plans, rfs, images):
            if rf.numel() != p._rf_numel or not rf.is_cuda or not rf.is_contiguous():
                raise _lib.FNumberError("fn_batch:InvalidRFData", "RF size does not match its plan")
            if img.numel() != p.info()["n_voxels"] * getattr(p, "n_frames", 1) or not img.is_contiguous():
                raise _lib.FNumberError("fn_batch:InvalidArgument", "image size does not match its plan")
        r = (C.c_void_p * n)(*[t.data_ptr() for t in rfs])
        d = (C.c_int32 * n)(*[_rf_dtype(t, "fn_batch") for t in rfs])
        im = (C.c_void_p * n)(*[t.data_ptr() for t in images])
        _lib.check(_lib.load().fn_batch_execute(self._h, r, d, im, _stream(stream, rfs[0].device)), "fn_batch")

    def info(self) -> dict:
        nodes, caps = C.c_int64(), C.c_int32()
        _lib.check(_lib.load().fn_batch_info(self._h, C.byref(nodes), C.byref(caps)), "fn_batch")
        ns, npl = C.c_int32(), C.c_int32()
        _lib.check(_lib.load().fn_batch_stacks(self._h, C.byref(ns), C.byref(npl)), "fn_batch")
        return {"graph_nodes": nodes.value, "captures": caps.value, "row_stacks": ns.value,
                "stacked_plans": npl.value}

    def close(self):
        if self._h:
            _lib.load().fn_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
