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
        if image.numel() != inf["n_voxels"]:
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


class PwPlan(_Plan):
    """Plan of das_pw for a fixed grid, acquisition and angle set."""
    prefix = "gpu_bf_das_pw_rf"

    def __init__(self, positions_x, positions_z, n_samples, n_elements, steering_angles, f_s, element_width,
                 element_pitch, c_0, f_bounds, index_t0=0.0, window=None, F_number=None, normalization=None,
                 index_gpu=0, semantics="cpu", precision="fp32", output_scale="mex"):
        super().__init__()
        window = window if window is not None else windows.tukey(0.2)
        F_number = F_number if F_number is not None else f_numbers.grating.angle_lb(60, 1.5)
        normalization = normalization if normalization is not None else normalizations.on()
        P, keep, info = pw_params(positions_x, positions_z, n_samples, n_elements, steering_angles, f_s,
                                  element_width, element_pitch, c_0, f_bounds, index_t0, window, F_number,
                                  normalization, index_gpu, semantics, precision, output_scale)
        _lib.check(_lib.load().fn_plan_pw_create(C.byref(P), C.byref(self._h)), self.prefix)
        self._keep = keep
        self.grid = info
        self._rf_numel = int(n_samples) * int(n_elements) * len(keep[2])


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
