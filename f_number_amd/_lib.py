"""ctypes binding of libfnumber_hip.so (include/fnumber_hip.h).

The library is built in-tree by ``f_number_amd.build``.  Loading fails loudly
when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# FN_LIB_PATH selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("FN_LIB_PATH") or os.path.join(_HERE, "libfnumber_hip.so")

# enums (include/fnumber_hip.h)
WINDOW = {"boxcar": 0, "hann": 1, "tukey": 2, "cosine": 3, "triangular": 4}
NORMALIZATION = {"off": 0, "on": 1, "windowed": 2}
SEMANTICS = {"cpu": 0, "mex": 1}
PRECISION = {"fp32": 0, "bf16x3": 1, "fp16x3": 2, "fp16": 3}
SCALE = {"mex": 0, "cpu": 1}
DTYPE = {"float32": 0, "float64": 1}
F_NUMBER = {"constant": 0, "grating.angle_lb": 1, "distance.constant": 2, "directivity.szabo": 3,
            "directivity.soft": 4, "directivity.rigid": 5, "directivity.perrot": 6, "distance.none": 7}

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class Window(C.Structure):
    _fields_ = [("id", C.c_int32), ("parameter", C.c_double)]


class FNumber(C.Structure):
    _fields_ = [("kind", C.c_int32), ("p0", C.c_double), ("p1", C.c_double), ("p2", C.c_double)]


class PwParams(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("positions_x", _dp), ("n_x", C.c_int32),
        ("positions_z", _dp), ("n_z", C.c_int32),
        ("n_samples", C.c_int32), ("n_elements", C.c_int32), ("n_angles", C.c_int32),
        ("steering_angles", _dp),
        ("f_s", C.c_double), ("element_width", C.c_double), ("element_pitch", C.c_double), ("c_0", C.c_double),
        ("f_lb", C.c_double), ("f_ub", C.c_double), ("index_t0", C.c_double),
        ("window", Window),
        ("f_number_values", _dp), ("n_f", C.c_int32),
        ("normalization", C.c_int32), ("normalization_window", Window),
        ("semantics", C.c_int32), ("precision", C.c_int32), ("output_scale", C.c_int32),
        ("index_gpu", C.c_int32), ("verbosity", C.c_int32), ("n_frames", C.c_int32),
        ("virtual_sources", _dp), ("line_scan", C.c_int32),
    ]


class SaParams(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("positions_x", _dp), ("n_x", C.c_int32),
        ("positions_z", _dp), ("n_z", C.c_int32),
        ("n_samples", C.c_int32), ("n_elements", C.c_int32),
        ("f_s", C.c_double), ("element_width", C.c_double), ("element_pitch", C.c_double), ("c_0", C.c_double),
        ("f_lb", C.c_double), ("f_ub", C.c_double), ("index_t0", C.c_double),
        ("window_tx", Window), ("window_rx", Window),
        ("f_number_tx_values", _dp), ("f_number_rx_values", _dp), ("n_f", C.c_int32),
        ("normalization_tx", C.c_int32), ("normalization_rx", C.c_int32),
        ("normalization_window_tx", Window), ("normalization_window_rx", Window),
        ("semantics", C.c_int32), ("precision", C.c_int32), ("output_scale", C.c_int32),
        ("index_gpu", C.c_int32), ("verbosity", C.c_int32),
    ]


class PlanInfo(C.Structure):
    _fields_ = [
        ("N_dft", C.c_int32), ("n_lb", C.c_int32), ("n_ub", C.c_int32), ("n_f", C.c_int32),
        ("n_tx", C.c_int32), ("n_voxels", C.c_int64), ("workspace_bytes", C.c_int64),
        ("flops_per_frame", C.c_double), ("useful_flops_per_frame", C.c_double),
        ("tile_voxels", C.c_int32), ("bin_split", C.c_int32), ("m_tiles", C.c_int32), ("k_steps", C.c_int32),
    ]


class SynthParams(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("seed", C.c_uint64),
        ("n_points", C.c_int32), ("n_speckle", C.c_int32), ("point_amplitude", C.c_double),
        ("x_min", C.c_double), ("x_max", C.c_double), ("z_min", C.c_double), ("z_max", C.c_double),
        ("n_samples", C.c_int32), ("n_elements", C.c_int32),
        ("f_s", C.c_double), ("element_pitch", C.c_double), ("c_0", C.c_double),
        ("f_c", C.c_double), ("fractional_bandwidth", C.c_double), ("snr_db", C.c_double),
        ("mode", C.c_int32), ("n_angles", C.c_int32), ("steering_angles", _dp),
        ("on_device", C.c_int32), ("index_gpu", C.c_int32),
        ("layer_thickness", C.c_double), ("layer_c", C.c_double), ("delay_offsets", _dp),
        ("tube_x", C.c_double), ("tube_z", C.c_double), ("tube_radius", C.c_double), ("tube_angle", C.c_double),
        ("tube_v_max", C.c_double), ("frame_time", C.c_double),
        ("virtual_sources", _dp),
    ]


class TuneResult(C.Structure):
    _fields_ = [("best_ms", C.c_double), ("default_ms", C.c_double), ("m_tiles", C.c_int32),
                ("bin_split", C.c_int32), ("n_candidates", C.c_int32)]


# every symbol include/fnumber_hip.h declares, with its prototype
PROTOTYPES = {
    "fn_abi_version": (C.c_int32, []),
    "fn_status_id": (C.c_char_p, [C.c_int32]),
    "fn_last_error": (C.c_char_p, []),
    "fn_last_report": (C.c_char_p, []),
    "fn_device_count": (C.c_int32, [C.POINTER(C.c_int32)]),
    "fn_pw_frequency_grid": (C.c_int32, [C.POINTER(PwParams)] + [C.POINTER(C.c_int32)] * 3),
    "fn_sa_frequency_grid": (C.c_int32, [C.POINTER(SaParams)] + [C.POINTER(C.c_int32)] * 3),
    "fn_element_pitch_over_lambda": (C.c_int32, [C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                                 C.c_int32, _dp]),
    "fn_f_number_values": (C.c_int32, [C.POINTER(FNumber), _dp, C.c_int32, _dp]),
    "fn_window_samples": (C.c_int32, [C.POINTER(Window), C.c_int32, _dp, C.c_int32, _dp]),
    "fn_pw_validate": (C.c_int32, [C.POINTER(PwParams)]),
    "fn_sa_validate": (C.c_int32, [C.POINTER(SaParams)]),
    "fn_das_pw": (C.c_int32, [C.POINTER(PwParams), C.c_void_p, C.c_int32, _fp, _fp]),
    "fn_das_sa": (C.c_int32, [C.POINTER(SaParams), C.c_void_p, C.c_int32, _fp, _fp, _fp]),
    "fn_plan_pw_create": (C.c_int32, [C.POINTER(PwParams), C.POINTER(C.c_void_p)]),
    "fn_plan_sa_create": (C.c_int32, [C.POINTER(SaParams), C.POINTER(C.c_void_p)]),
    "fn_plan_info_get": (C.c_int32, [C.c_void_p, C.POINTER(PlanInfo)]),
    "fn_plan_execute": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "fn_plan_set_profiling": (C.c_int32, [C.c_void_p, C.c_int32]),
    "fn_plan_get_profile": (C.c_int32, [C.c_void_p, _dp, _dp, C.POINTER(C.c_int32)]),
    "fn_plan_destroy": (C.c_int32, [C.c_void_p]),
    "fn_plan_autotune": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                                     C.POINTER(TuneResult), C.c_void_p]),
    "fn_plan_set_tiling": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32]),
    "fn_batch_create": (C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "fn_batch_execute": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_void_p), C.c_void_p]),
    "fn_batch_info": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "fn_batch_set_persistent": (C.c_int32, [C.c_void_p, C.c_int32]),
    "fn_batch_set_stacking": (C.c_int32, [C.c_void_p, C.c_int32]),
    "fn_batch_stacks": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "fn_batch_destroy": (C.c_int32, [C.c_void_p]),
    "fn_plan_focused_signal": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fn_plan_set_tof": (C.c_int32, [C.c_void_p, C.c_double, _dp]),
    "fn_plan_execute_tangent": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "fn_plan_adjoint": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "fn_plan_vjp": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    "fn_das_pw_adjoint": (C.c_int32, [C.POINTER(PwParams), C.c_double, _dp, _fp, _fp]),
    "fn_das_pw_noalloc": (C.c_int32, [C.POINTER(PwParams), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "fn_synth_rf": (C.c_int32, [C.POINTER(SynthParams), C.c_void_p]),
    "fn_synth_scatterers": (C.c_int32, [C.POINTER(SynthParams), _dp, _dp, _dp]),
}

_lib = None


def load():
    """Load libfnumber_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m f_number_amd.build` "
                "(the HIP library is required; there is no fallback path)")
        # One HIP runtime per process: torch ships its own libamdhip64 (SONAME
        # libamdhip64.so.7, but its modules link the unversioned name).  Loaded
        # first, it satisfies this library's libamdhip64.so.7 dependency; loaded
        # after /opt/rocm's copy it would be a second runtime that sees no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.fn_abi_version() != 1:
            raise RuntimeError("libfnumber_hip.so ABI version mismatch")
        _lib = L
    return _lib


class FNumberError(Exception):
    """Error with the reference's MATLAB identifier (e.g. 'gpu_bf_das_pw_rf:InvalidRFData')."""

    def __init__(self, identifier: str, message: str):
        super().__init__(f"{identifier}: {message}")
        self.identifier = identifier
        self.message = message


def check(status: int, prefix: str) -> None:
    if status != 0:
        L = load()
        ident = L.fn_status_id(status).decode()
        raise FNumberError(f"{prefix}:{ident}", L.fn_last_error().decode())
