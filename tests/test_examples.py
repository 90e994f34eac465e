"""Experimental-data I/O (data_RF.mat, examples.m:38) and the examples pipeline
(examples.m:56-249).  The reference's data_RF.mat is missing (.MISSING_LARGE_BLOBS),
so a seeded synthetic acquisition is written in the same format."""
import os

import numpy as np
import pytest

from conftest import rel_l2


def _write(tmp_path, n_angles=3, n_t=768, n_el=32):
    from f_number_amd import examples, synth

    w = synth.scaled(synth.cpwc_small(), 16, 16, n_angles=n_angles, n_samples=n_t, n_speckle=400, n_elements=n_el)
    rf = np.transpose(w.rf_host(), (2, 1, 0))  # MATLAB [N_t, N_el, N_angles]
    path = os.path.join(tmp_path, "data_RF.mat")
    examples.save_data_rf(path, rf, w.f_s, w.f_bounds[0], w.f_bounds[1], w.angles_deg, w.width, w.pitch, w.c_0)
    return path, w, rf


def test_data_rf_round_trip(tmp_path):
    from f_number_amd import examples

    path, w, rf = _write(tmp_path)
    d = examples.load_data_rf(path)
    assert d["data_RF"].dtype == np.float32 and np.array_equal(d["data_RF"], rf)
    assert d["f_s"] == w.f_s and d["c_avg"] == w.c_0 and d["element_pitch"] == w.pitch
    assert np.array_equal(d["steering_angles_deg"], w.angles_deg)


def test_data_rf_errors(tmp_path):
    import scipy.io

    from f_number_amd import FNumberError, examples

    p = os.path.join(tmp_path, "bad.mat")
    scipy.io.savemat(p, {"data_RF": np.zeros((8, 4, 2)), "f_s": 1.0})
    with pytest.raises(FNumberError, match="MissingVariables"):
        examples.load_data_rf(p)
    path, w, rf = _write(tmp_path)
    m = scipy.io.loadmat(path)
    m["steering_angles_deg"] = np.zeros((1, 5))
    scipy.io.savemat(p, {k: v for k, v in m.items() if not k.startswith("__")})
    with pytest.raises(FNumberError, match="InvalidSteeringAngles"):
        examples.load_data_rf(p)


def test_grid_and_db():
    from f_number_amd import examples

    x, z = examples.grid(0.3048e-3)
    assert x.size == 512 and z.size == 512
    assert np.isclose(x[0], -255.5 * 0.3048e-3 / 4) and np.isclose(z[0], 64 * 0.3048e-3 / 4)
    db = examples.to_db(np.array([[1.0, 0.1, 1e-9]]), 60)
    assert np.allclose(db, [[0.0, -20.0, -60.0]])


@pytest.mark.gpu
def test_examples_pipeline_on_gpu(tmp_path, oracle):
    from f_number_amd import examples

    path, w, rf = _write(tmp_path)
    d = examples.load_data_rf(path)
    print("case examples pipeline", flush=True)
    res = examples.run_examples(d, n_x=24, n_z=20, output_scale="cpu")
    x, z = res["positions_x"], res["positions_z"]
    s1 = res[1]
    per_angle = [v for k, v in s1.items() if k.startswith("angle_")]
    assert len(per_angle) == 3 and rel_l2(s1["compound"], np.sum(per_angle, axis=0)) < 1e-12
    assert [len(res[k]) for k in (2, 3, 4, 5)] == [3, 4, 2, 4]
    # the first angle's default image against the FP64 oracle (examples.m:68 defaults:
    # Tukey(0.2), angle_lb(60, 1.5), normalization on, index_t0 = 8)
    ref, _ = oracle.das_pw(x, z, np.transpose(rf, (2, 1, 0))[:1], d["f_s"], np.deg2rad(d["steering_angles_deg"][:1]),
                           d["element_width"], d["element_pitch"], d["c_avg"], (d["f_lb"], d["f_ub"]), 8,
                           ("tukey", 0.2), oracle.fnumber("angle_lb", 60, 1.5, 10), "on")
    e = rel_l2(per_angle[0], ref)
    print(f"  rel_l2 {e:.2e}", flush=True)
    assert e <= 1e-4
    for sec in (2, 3, 4, 5):
        for img in res[sec].values():
            assert img.shape == (z.size, x.size) and np.all(np.isfinite(img))
