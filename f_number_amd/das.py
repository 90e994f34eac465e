"""das_pw / das_sa: host-side mirror of the reference's MATLAB entry points.

Argument checks follow das_pw.m:85-290 and das_sa.m:90-280 (same order and
error identifiers); the computation is the GPU branch (das_pw.m:296-300 ->
gpu_bf_das_pw_rf, das_sa.m:283-287 -> gpu_bf_saft_rf), served by
libfnumber_hip.so.  The reference's CPU branch is not shipped.

Layouts follow MATLAB: data_RF is [N_t, N_el] (or [N_t, N_el, N_angles] for
compounding several steering angles in one call) for das_pw and
[N_t, N_rx, N_tx] for das_sa; images are complex [N_z, N_x].
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Optional

import numpy as np

from . import _lib, f_numbers, normalizations, platforms, windows
from ._lib import FNumberError

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def _err(ident, msg):
    raise FNumberError(ident, msg)


def _row(a, name, prefix):
    a = np.asarray(a)
    if not (np.issubdtype(a.dtype, np.number) and np.isrealobj(a) and (a.ndim == 1 or (a.ndim == 2 and a.shape[0] == 1))):
        _err(f"{prefix}:NoNumericAndRealRowVector", f"{name} must be a numeric and real-valued row vector!")
    return np.ascontiguousarray(a.reshape(-1), dtype=np.float64)


def _scalar(v, name, prefix):
    a = np.asarray(v)
    if not (np.issubdtype(a.dtype, np.number) and np.isrealobj(a) and a.size == 1):
        _err(f"{prefix}:NoNumericAndRealScalar", f"{name} must be a numeric and real-valued scalar!")
    return float(a.reshape(-1)[0])


def _common_checks(prefix, positions_x, positions_z, data_RF, f_s, element_width, element_pitch, c_0,
                   f_bounds, index_t0, rf_ndim_ok):
    px = _row(positions_x, "positions_x", prefix)
    if not np.all(np.isfinite(px)):
        _err(f"{prefix}:NoFiniteXCoordinates", "positions_x must contain finite x-coordinates!")
    pz = _row(positions_z, "positions_z", prefix)
    if not (np.all(pz > 0) and np.all(np.isfinite(pz))):
        _err(f"{prefix}:NoPositiveAndFiniteZCoordinates", "positions_z must contain positive and finite z-coordinates!")
    rf = np.asarray(data_RF)
    if not (np.issubdtype(rf.dtype, np.number) and np.isrealobj(rf) and rf.ndim in rf_ndim_ok):
        _err(f"{prefix}:NoNumericAndRealMatrix", "data_RF must be a numeric and real-valued matrix!")
    if not (rf.shape[0] > 1 and rf.shape[1] > 1):
        _err(f"{prefix}:NoSamples", "data_RF must contain at least two temporal and two spatial samples!")
    if not np.all(np.isfinite(rf)):
        _err(f"{prefix}:NoFiniteRFData", "data_RF must contain finite samples!")
    f_s = _scalar(f_s, "f_s", prefix)
    if not (f_s > 0 and math.isfinite(f_s)):
        _err(f"{prefix}:InvalidSamplingRate", "f_s must be positive and finite!")
    w = _scalar(element_width, "element_width", prefix)
    if not (w > 0 and math.isfinite(w)):
        _err(f"{prefix}:InvalidElementWidth", "element_width must be positive and finite!")
    p = _scalar(element_pitch, "element_pitch", prefix)
    if not (p > w and math.isfinite(p)):
        _err(f"{prefix}:InvalidElementPitch", "element_pitch must be larger than element_width and finite!")
    c = _scalar(c_0, "c_0", prefix)
    if not (c > 0 and math.isfinite(c)):
        _err(f"{prefix}:InvalidSpeedOfSound", "c_0 must be positive and finite!")
    if f_bounds is None or np.size(f_bounds) == 0:
        f_bounds = [np.finfo(float).eps, f_s / 2]
    fb = _row(f_bounds, "f_bounds", prefix)
    if not (fb.size == 2 and fb[0] > 0 and fb[1] > fb[0] and fb[1] <= f_s / 2):
        _err(f"{prefix}:InvalidFrequencyBounds", "f_bounds must be a strictly monotonic increasing sequence!")
    if index_t0 is None or np.size(index_t0) == 0:
        index_t0 = 0
    t0 = _scalar(index_t0, "index_t0", prefix)
    if not (t0 >= 0 and t0 == math.floor(t0)):
        _err(f"{prefix}:InvalidTimeIndex", "index_t0 must be a nonnegative integer!")
    return px, pz, rf, f_s, w, p, c, fb, t0


def _window(win, default, ident, prefix):
    if win is None:
        win = default
    if not isinstance(win, windows.window):
        _err(f"{prefix}:{ident}", "window must be scalar windows.window!")
    return win


def _fnumber(F, ident, prefix):
    if F is None:
        F = f_numbers.grating.angle_lb(60, 1.5)
    if not isinstance(F, f_numbers.f_number):
        _err(f"{prefix}:{ident}", "F_number must be scalar f_numbers.f_number!")
    return F


def _norm(n, ident, prefix):
    if n is None:
        n = normalizations.on()
    if not isinstance(n, normalizations.normalization):
        _err(f"{prefix}:{ident}", "normalization must be scalar normalizations.normalization!")
    return n


def _platform(pf, prefix):
    if pf is None:
        pf = platforms.gpu(0)
    if not isinstance(pf, platforms.platform):
        _err(f"{prefix}:NoPlatform", "platform must be scalar platforms.platform!")
    if not isinstance(pf, platforms.gpu):
        _err(f"{prefix}:UnsupportedPlatform",
             "only platforms.gpu is served; the MATLAB CPU branch is not part of this package")
    return pf


def _image_out(buf, n_x, n_z):
    return (buf[0::2] + 1j * buf[1::2]).astype(np.complex64).reshape(n_x, n_z).T


def pw_params(px, pz, n_t, n_el, angles, f_s, w, p, c, fb, t0, window, F, normalization, index_gpu=0,
              semantics="cpu", precision="fp32", output_scale="mex", verbosity=0, n_frames=1,
              virtual_sources=None, line_scan=False):
    """Build the fn_pw_params struct (and the arrays it points to) for libfnumber_hip.
    virtual_sources: (N_a, 2) array of (x_vs, z_vs) in m replacing the steering angles.
    line_scan: grid column i is formed from transmit i alone (N_a == N_x)."""
    L = _lib.load()
    px = np.ascontiguousarray(px, dtype=np.float64)
    pz = np.ascontiguousarray(pz, dtype=np.float64)
    angles = np.ascontiguousarray(np.atleast_1d(angles), dtype=np.float64)
    P = _lib.PwParams()
    P.struct_size = C.sizeof(_lib.PwParams)
    P.positions_x, P.n_x = px.ctypes.data_as(_dp), px.size
    P.positions_z, P.n_z = pz.ctypes.data_as(_dp), pz.size
    P.n_samples, P.n_elements, P.n_angles = int(n_t), int(n_el), angles.size
    P.steering_angles = angles.ctypes.data_as(_dp)
    P.f_s, P.element_width, P.element_pitch, P.c_0 = f_s, w, p, c
    P.f_lb, P.f_ub = float(fb[0]), float(fb[1])
    P.index_t0 = t0
    P.window = window._c()
    P.normalization = normalization.code
    P.normalization_window = normalization.window._c()
    P.semantics = _lib.SEMANTICS[semantics]
    P.precision = _lib.PRECISION[precision]
    P.output_scale = _lib.SCALE[output_scale]
    P.index_gpu = int(index_gpu)
    P.verbosity = int(verbosity)
    P.n_frames = int(n_frames)
    P.line_scan = int(bool(line_scan))
    vs = None
    if virtual_sources is not None:
        vs = np.ascontiguousarray(np.asarray(virtual_sources, dtype=np.float64).reshape(-1, 2))
        if vs.shape[0] != angles.size:
            _err("gpu_bf_das_pw_rf:InvalidSteeringAngle", "one virtual source (x_vs, z_vs) per transmit is required")
        P.virtual_sources = vs.ctypes.data_as(_dp)
    N, lb, ub = C.c_int32(), C.c_int32(), C.c_int32()
    _lib.check(L.fn_pw_frequency_grid(C.byref(P), C.byref(N), C.byref(lb), C.byref(ub)), "gpu_bf_das_pw_rf")
    r = np.empty(ub.value - lb.value + 1)
    _lib.check(L.fn_element_pitch_over_lambda(p, f_s, c, N.value, lb.value, ub.value, r.ctypes.data_as(_dp)),
               "gpu_bf_das_pw_rf")
    Fv = np.ascontiguousarray(F.compute_values(r), dtype=np.float64)
    P.f_number_values, P.n_f = Fv.ctypes.data_as(_dp), Fv.size
    keep = (px, pz, angles, Fv, vs)
    return P, keep, dict(N_dft=N.value, n_lb=lb.value, n_ub=ub.value, F_number_values=Fv)


def sa_params(px, pz, n_t, n_el, f_s, w, p, c, fb, t0, window_tx, window_rx, F_tx, F_rx, norm_tx, norm_rx,
              index_gpu=0, semantics="cpu", precision="fp32", output_scale="mex", verbosity=0):
    """Build the fn_sa_params struct for libfnumber_hip."""
    L = _lib.load()
    px = np.ascontiguousarray(px, dtype=np.float64)
    pz = np.ascontiguousarray(pz, dtype=np.float64)
    P = _lib.SaParams()
    P.struct_size = C.sizeof(_lib.SaParams)
    P.positions_x, P.n_x = px.ctypes.data_as(_dp), px.size
    P.positions_z, P.n_z = pz.ctypes.data_as(_dp), pz.size
    P.n_samples, P.n_elements = int(n_t), int(n_el)
    P.f_s, P.element_width, P.element_pitch, P.c_0 = f_s, w, p, c
    P.f_lb, P.f_ub = float(fb[0]), float(fb[1])
    P.index_t0 = t0
    P.window_tx, P.window_rx = window_tx._c(), window_rx._c()
    P.normalization_tx, P.normalization_rx = norm_tx.code, norm_rx.code
    P.normalization_window_tx, P.normalization_window_rx = norm_tx.window._c(), norm_rx.window._c()
    P.semantics = _lib.SEMANTICS[semantics]
    P.precision = _lib.PRECISION[precision]
    P.output_scale = _lib.SCALE[output_scale]
    P.index_gpu = int(index_gpu)
    P.verbosity = int(verbosity)
    N, lb, ub = C.c_int32(), C.c_int32(), C.c_int32()
    _lib.check(L.fn_sa_frequency_grid(C.byref(P), C.byref(N), C.byref(lb), C.byref(ub)), "gpu_bf_saft_rf")
    r = np.empty(ub.value - lb.value + 1)
    _lib.check(L.fn_element_pitch_over_lambda(p, f_s, c, N.value, lb.value, ub.value, r.ctypes.data_as(_dp)),
               "gpu_bf_saft_rf")
    Ftx = np.ascontiguousarray(F_tx.compute_values(r), dtype=np.float64)
    Frx = np.ascontiguousarray(F_rx.compute_values(r), dtype=np.float64)
    P.f_number_tx_values, P.f_number_rx_values, P.n_f = Ftx.ctypes.data_as(_dp), Frx.ctypes.data_as(_dp), Ftx.size
    keep = (px, pz, Ftx, Frx)
    return P, keep, dict(N_dft=N.value, n_lb=lb.value, n_ub=ub.value, F_number_tx_values=Ftx, F_number_rx_values=Frx)


def das_pw(positions_x, positions_z, data_RF, f_s, steering_angle, element_width, element_pitch, c_0,
           f_bounds=None, index_t0=None, window=None, F_number=None, normalization=None, platform=None, *,
           semantics: str = "cpu", precision: str = "fp32", output_scale: str = "mex",
           return_weights: bool = False, verbosity: int = 0, nargout: int = 1, virtual_sources=None):
    """Fourier-domain DAS of steered plane waves (das_pw.m:1).

    Returns the complex64 image [N_z, N_x] in the MEX scale 2/N_dft * sum
    (gpu_bf_das_pw_rf.cu:1118-1119) unless output_scale='cpu' (2 * sum,
    das_pw.m:384,467).  steering_angle may hold N_a angles with data_RF
    [N_t, N_el, N_a]: the result is the compounded image.  With
    return_weights=True also returns the summed receive apodization weights of
    the last frequency (the MEX's second output) and an info dict.
    nargout=2 or 3 returns (image, F_number_values[, signal]) like the CPU path
    (das_pw.m:1): the F-numbers per bin and depth [N_f, N_z] (clamped at
    sqrt(f z / 2.88 c) in CPU semantics, das_pw.m:374-377) and the focused RF
    signal of the last voxel [N_dft] (das_pw.m:477-481; angles summed).
    virtual_sources: (N_a, 2) (x_vs, z_vs) in m per transmit instead of the plane
    waves of steering_angle (which then only gives N_a): diverging (z_vs < 0) or
    focused (z_vs > 0) waves (fnumber_hip.h fn_pw_params.virtual_sources).
    """
    prefix = "das_pw"
    px, pz, rf, f_s, w, p, c, fb, t0 = _common_checks(prefix, positions_x, positions_z, data_RF, f_s,
                                                      element_width, element_pitch, c_0, f_bounds, index_t0, (2, 3))
    ang = np.atleast_1d(np.asarray(steering_angle, dtype=np.float64)).reshape(-1)
    if not (np.isrealobj(ang) and ang.size >= 1):
        _err(f"{prefix}:NoNumericAndRealScalar", "steering_angle must be a numeric and real-valued scalar!")
    if not np.all((ang > -math.pi / 2) & (ang < math.pi / 2)):
        _err(f"{prefix}:InvalidSteeringAngle",
             "steering_angle must be larger than negative pi over two and smaller than pi over two!")
    n_a = rf.shape[2] if rf.ndim == 3 else 1
    if ang.size != n_a:
        _err(f"{prefix}:NoNumericAndRealScalar", "steering_angle must hold one angle per RF data set!")
    win = _window(window, windows.tukey(0.2), "NoWindow", prefix)
    F = _fnumber(F_number, "NoFNumber", prefix)
    nrm = _norm(normalization, "NoNormalization", prefix)
    pf = _platform(platform, prefix)
    n_t, n_el = rf.shape[0], rf.shape[1]
    P, keep, info = pw_params(px, pz, n_t, n_el, ang, f_s, w, p, c, fb, t0, win, F, nrm, pf.index,
                              semantics, precision, output_scale, verbosity, virtual_sources=virtual_sources)
    if nargout >= 2:
        f = f_s * np.arange(info["n_lb"], info["n_ub"] + 1) / info["N_dft"]
        Fz = np.repeat(info["F_number_values"][:, None], pz.size, axis=1)
        if semantics == "cpu":
            Fz = np.minimum(Fz, np.sqrt(f[:, None] * pz[None, :] / (2.88 * c)))
        outs = (Fz,)
        if nargout >= 3:
            from .plan import PwPlan

            plan = PwPlan(px, pz, n_t, n_el, ang, f_s, w, p, c, fb, t0, window=win, F_number=F, normalization=nrm,
                          index_gpu=pf.index, semantics=semantics, precision=precision, output_scale=output_scale,
                          virtual_sources=virtual_sources)
            image, sig = _execute_with_signal(plan, rf, px.size, pz.size, pf.index)
            return (image,) + outs + (sig,)
    rfF = np.asfortranarray(rf, dtype=np.float64 if rf.dtype == np.float64 else np.float32)
    dtype = _lib.DTYPE["float64" if rfF.dtype == np.float64 else "float32"]
    img = np.empty(2 * px.size * pz.size, dtype=np.float32)
    wts = np.empty(px.size * pz.size, dtype=np.float32) if return_weights else None
    L = _lib.load()
    _lib.check(L.fn_das_pw(C.byref(P), rfF.ctypes.data, dtype, img.ctypes.data_as(_fp),
                           wts.ctypes.data_as(_fp) if wts is not None else None), "gpu_bf_das_pw_rf")
    del keep
    image = _image_out(img, px.size, pz.size)
    if nargout >= 2:
        return (image,) + outs
    if return_weights:
        return image, wts.reshape(px.size, pz.size).T, info
    return image


def _execute_with_signal(plan, rf, n_x, n_z, index_gpu):
    """image [N_z, N_x] and focused signal [N_dft] through a plan (device path)"""
    import torch

    dev = torch.device("cuda", index_gpu)
    # MATLAB [N_t, N_el, N_tx] -> device (N_tx, N_el, N_t)
    host = np.ascontiguousarray(np.transpose(rf if rf.ndim == 3 else rf[:, :, None], (2, 1, 0)))
    x = torch.from_numpy(host.astype(np.float64 if rf.dtype == np.float64 else np.float32)).to(dev)
    img = torch.empty((n_x, n_z), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)
    plan.execute(x, img, stream=stream)
    sig = plan.focused_signal(stream=stream, device=dev)
    torch.cuda.synchronize(dev)
    out = img.cpu().numpy().T.copy(), sig.cpu().numpy()
    plan.close()
    return out


def das_sa(positions_x, positions_z, data_RF, f_s, element_width, element_pitch, c_0, f_bounds=None,
           index_t0=None, window_tx=None, window_rx=None, F_number_tx=None, F_number_rx=None,
           normalization_tx=None, normalization_rx=None, platform=None, *, semantics: str = "cpu",
           precision: str = "fp32", output_scale: str = "mex", return_weights: bool = False, verbosity: int = 0,
           nargout: int = 1):
    """Fourier-domain DAS of a full synthetic-aperture dataset (das_sa.m:1).

    data_RF is [N_t, N_rx, N_tx] with N_rx == N_tx.  The receive window,
    F-number and normalization default to the transmit ones.  Returns the
    complex64 image [N_z, N_x] (MEX scale unless output_scale='cpu'); with
    return_weights=True also the transmit and receive weight sums of the last
    frequency (gpu_bf_saft_rf.cu outputs 2 and 3) and an info dict.
    nargout=2..4 returns (image, F_number_tx_values, F_number_rx_values[, signal])
    like the CPU path (das_sa.m:1; F per bin [N_f], no depth clamp, das_sa.m:354-358).
    """
    prefix = "das_sa"
    px, pz, rf, f_s, w, p, c, fb, t0 = _common_checks(prefix, positions_x, positions_z, data_RF, f_s,
                                                      element_width, element_pitch, c_0, f_bounds, index_t0, (3,))
    if rf.shape[1] != rf.shape[2]:
        _err("gpu_bf_saft_rf:InvalidRFData", "Size of RF data must be equal along 2nd and 3rd dimensions!")
    wtx = _window(window_tx, windows.tukey(0.2), "NoWindowTX", prefix)
    wrx = _window(window_rx, wtx, "NoWindowRX", prefix)
    Ftx = _fnumber(F_number_tx, "NoFNumberTX", prefix)
    Frx = _fnumber(F_number_rx if F_number_rx is not None else Ftx, "NoFNumberRX", prefix)
    ntx = _norm(normalization_tx, "NoNormalizationTX", prefix)
    nrx = _norm(normalization_rx if normalization_rx is not None else ntx, "NoNormalizationRX", prefix)
    pf = _platform(platform, prefix)
    n_t, n_el = rf.shape[0], rf.shape[1]
    P, keep, info = sa_params(px, pz, n_t, n_el, f_s, w, p, c, fb, t0, wtx, wrx, Ftx, Frx, ntx, nrx, pf.index,
                              semantics, precision, output_scale, verbosity)
    if nargout >= 2:
        outs = (info["F_number_tx_values"].copy(), info["F_number_rx_values"].copy())
        if nargout >= 4:
            from .plan import SaPlan

            plan = SaPlan(px, pz, n_t, n_el, f_s, w, p, c, fb, t0, window_tx=wtx, window_rx=wrx, F_number_tx=Ftx,
                          F_number_rx=Frx, normalization_tx=ntx, normalization_rx=nrx, index_gpu=pf.index,
                          semantics=semantics, precision=precision, output_scale=output_scale)
            image, sig = _execute_with_signal(plan, rf, px.size, pz.size, pf.index)
            return (image,) + outs + (sig,)
    rfF = np.asfortranarray(rf, dtype=np.float64 if rf.dtype == np.float64 else np.float32)
    dtype = _lib.DTYPE["float64" if rfF.dtype == np.float64 else "float32"]
    img = np.empty(2 * px.size * pz.size, dtype=np.float32)
    wt = np.empty(px.size * pz.size, dtype=np.float32) if return_weights else None
    wr = np.empty(px.size * pz.size, dtype=np.float32) if return_weights else None
    L = _lib.load()
    _lib.check(L.fn_das_sa(C.byref(P), rfF.ctypes.data, dtype, img.ctypes.data_as(_fp),
                           wt.ctypes.data_as(_fp) if wt is not None else None,
                           wr.ctypes.data_as(_fp) if wr is not None else None), "gpu_bf_saft_rf")
    del keep
    image = _image_out(img, px.size, pz.size)
    if nargout >= 2:
        return (image,) + outs
    if return_weights:
        return image, wt.reshape(px.size, pz.size).T, wr.reshape(px.size, pz.size).T, info
    return image
