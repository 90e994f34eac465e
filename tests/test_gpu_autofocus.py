"""Autofocus (differentiation form) on the layered-medium acquisition at a reduced
grid: gradient ascent on the image intensity from c = 1480 m/s must raise the
objective, move the speed of sound towards the medium's effective value
(6 mm at 1450 m/s over 1540 m/s: about 1505-1535 m/s at 10-40 mm depth) and
shrink the per-element delay error (mean removed) of the N(0, 20 ns) offsets."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_autofocus_layer_recovers_speed_and_delays():
    import torch

    import f_number_amd as fa
    from f_number_amd import autofocus as AF
    from f_number_amd import synth

    fa.load()
    w = synth.autofocus_layer()
    w.positions_x = np.linspace(w.positions_x[0], w.positions_x[-1], 128)
    w.positions_z = np.linspace(w.positions_z[0], w.positions_z[-1], 128)
    rf = w.rf_device("cuda:0")
    kw = dict(window=fa.windows.tukey(0.5), F_number=fa.f_numbers.grating.angle_lb(60, 1.5),
              normalization=fa.normalizations.on())
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     1480.0, w.f_bounds, 0.0, **kw)
    res = AF.autofocus(plan, rf, n_iter=40, c0=1480.0)
    plan.close()
    h = res["history"]
    truth = w.delay_offsets - w.delay_offsets.mean()
    d = res["delays"] - res["delays"].mean()
    e0, e1 = np.std(truth), np.std(d - truth)
    print(f"J {h[0]['J']:.4e} -> {h[-1]['J']:.4e}; c {res['c']:.1f}; delay error {e0 * 1e9:.1f} -> {e1 * 1e9:.1f} ns; "
          f"{res['seconds']:.2f} s", flush=True)
    assert h[-1]["J"] > 1.5 * h[0]["J"]
    assert 1500.0 <= res["c"] <= 1545.0
    assert e1 < 0.6 * e0
