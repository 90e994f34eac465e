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
