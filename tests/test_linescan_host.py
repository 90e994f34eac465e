"""Line-by-line scanning: argument checks that fail before any device work."""
import numpy as np
import pytest

import f_number_amd as fa
from f_number_amd import synth


def _args(w):
    return (w.n_samples, w.n_elements, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds)


@pytest.mark.parametrize("focus_z", [0.0, -5e-3, np.inf, np.nan])
def test_focus_depth_must_lie_below_the_array(focus_z):
    w = synth.line_scan(8, 16)
    with pytest.raises(fa.FNumberError, match="das_lbl:InvalidFocus"):
        fa.LineScan(w.positions_x, w.positions_z, *_args(w), focus_z)


def test_one_focus_per_line():
    w = synth.line_scan(8, 16)
    with pytest.raises(fa.FNumberError, match="das_lbl:InvalidFocus"):
        fa.LineScan(w.positions_x, w.positions_z, *_args(w), 20e-3, focus_x=np.zeros(7))


def test_one_transmit_event_per_line():
    w = synth.line_scan(8, 16)
    rf = np.zeros((w.n_samples, w.n_elements, 7), dtype=np.float32)
    with pytest.raises(fa.FNumberError, match="das_lbl:InvalidRFData"):
        fa.das_lbl(w.positions_x, w.positions_z, rf, w.f_s, 20e-3, w.width, w.pitch, w.c_0)


def test_rf_must_be_real_float():
    w = synth.line_scan(8, 16)
    rf = np.zeros((w.n_samples, w.n_elements, 8), dtype=np.complex64)
    with pytest.raises(fa.FNumberError, match="das_lbl:InvalidRFData"):
        fa.das_lbl(w.positions_x, w.positions_z, rf, w.f_s, 20e-3, w.width, w.pitch, w.c_0)


def test_line_scan_workload_focuses_each_line():
    w = synth.line_scan(8, 16, focus_z=25e-3)
    assert w.n_tx == 8 and w.virtual_sources.shape == (8, 2)
    assert np.array_equal(w.virtual_sources[:, 0], w.positions_x) and np.all(w.virtual_sources[:, 1] == 25e-3)


def _pw(**kw):
    from f_number_amd.das import pw_params

    base = dict(px=np.linspace(-1e-3, 1e-3, 4), pz=np.linspace(5e-3, 8e-3, 5), n_t=256, n_el=16,
                angles=np.zeros(4), f_s=20.832e6, w=0.27e-3, p=0.3048e-3, c=1540.0, fb=(2e6, 7e6), t0=0.0,
                window=fa.windows.hann(), F=fa.f_numbers.constant(1.5), normalization=fa.normalizations.on(),
                virtual_sources=np.stack([np.linspace(-1e-3, 1e-3, 4), np.full(4, 6e-3)], axis=1), line_scan=True)
    base.update(kw)
    return pw_params(**base)


def _validate(P):
    import ctypes as C

    from f_number_amd import _lib

    L = _lib.load()
    st = L.fn_pw_validate(C.byref(P))
    return L.fn_status_id(st).decode(), L.fn_last_error().decode() if st else ""


def test_line_scan_params_validate():
    P, keep, info = _pw()
    assert _validate(P) == ("OK", "")


@pytest.mark.parametrize("kw,ident,msg", [
    (dict(angles=np.zeros(3), virtual_sources=np.zeros((3, 2)) + [0.0, 6e-3]), "InvalidArgument",
     "one transmit per grid column"),
    (dict(precision="fp16"), "Unsupported", "FP32 only"),
    (dict(n_frames=2), "Unsupported", "one frame"),
    (dict(normalization=fa.normalizations.windowed(fa.windows.hann())), "Unsupported", "windowed"),
])
def test_line_scan_params_refused(kw, ident, msg):
    P, keep, info = _pw(**kw)
    got, text = _validate(P)
    assert got == ident and msg in text


def test_line_scan_flag_is_zero_or_one():
    P, keep, info = _pw()
    P.line_scan = 2
    got, text = _validate(P)
    assert got == "InvalidArgument" and "line_scan" in text


def test_line_scan_platform_must_be_a_gpu():
    w = synth.line_scan(8, 16)
    rf = np.zeros((w.n_samples, w.n_elements, 8), dtype=np.float32)
    with pytest.raises(fa.FNumberError, match="das_lbl:UnsupportedPlatform"):
        fa.das_lbl(w.positions_x, w.positions_z, rf, w.f_s, 20e-3, w.width, w.pitch, w.c_0,
                   platform=fa.platforms.cpu())
    with pytest.raises(fa.FNumberError, match="das_lbl:NoPlatform"):
        fa.das_lbl(w.positions_x, w.positions_z, rf, w.f_s, 20e-3, w.width, w.pitch, w.c_0, platform="gpu")
