"""Window samples (+windows/<name>.m for CPU semantics, gpu_bf_windows.cu for MEX)."""
import math

import numpy as np
import pytest

import f_number_amd as fa


def test_closed_form_samples(oracle):
    t = np.array([0.0, 0.25, 0.5, 0.75, 0.999, 1.0, 1.5, 2.0])
    box = oracle.window("boxcar", 0, t)
    assert list(box) == [1, 1, 1, 1, 1, 0, 0, 0]
    hann = oracle.window("hann", 0, t)
    assert hann[0] == 1.0 and hann[2] == pytest.approx(0.5) and hann[5] == pytest.approx(0.0, abs=1e-16)
    assert hann[7] == pytest.approx(1.0)  # CPU hann is periodic, not truncated (hann.m:62)
    tri = oracle.window("triangular", 0, t)
    assert tri[7] == pytest.approx(-1.0)  # CPU triangular is not truncated (triangular.m:62)
    tk = oracle.window("tukey", 0.5, t)
    assert list(tk[:3]) == [1.0, 1.0, 1.0]  # flat for |t| <= 1 - alpha
    assert tk[3] == pytest.approx(0.5)      # mid-taper: (1 + cos(pi/2)) / 2
    assert list(tk[5:]) == [0.0, 0.0, 0.0]


def test_symmetry_and_tukey_limits(oracle):
    t = np.linspace(-1.2, 1.2, 97)
    for name, p in [("hann", 0), ("tukey", 0.3), ("triangular", 0), ("boxcar", 0)]:
        assert np.array_equal(oracle.window(name, p, t), oracle.window(name, p, -t))
    # tukey -> hann as alpha -> 1 on |t| < 1
    tt = np.linspace(0, 0.99, 50)
    assert np.allclose(oracle.window("tukey", 0.999999, tt), oracle.window("hann", 0, tt), atol=1e-5)


def test_mex_truncation(oracle):
    t = np.array([0.5, 1.0, 1.5, 2.0])
    for name in ["hann", "triangular", "tukey", "boxcar"]:
        w = oracle.window(name, 0.5, t, "mex")
        assert list(w[1:]) == [0.0, 0.0, 0.0]
    assert oracle.window("hann", 0, [0.5], "mex")[0] == pytest.approx(0.5)


@pytest.mark.parametrize("w,name,p", [(fa.windows.boxcar(), "boxcar", 0), (fa.windows.hann(), "hann", 0),
                                      (fa.windows.tukey(0.2), "tukey", 0.2), (fa.windows.tukey(0.5), "tukey", 0.5),
                                      (fa.windows.triangular(), "triangular", 0)])
@pytest.mark.parametrize("sem", ["cpu", "mex"])
def test_product_matches_oracle(oracle, w, name, p, sem):
    t = np.linspace(-2.5, 2.5, 1001)
    assert np.array_equal(w.compute_samples(t, sem), oracle.window(name, p, t, sem)) or \
        np.allclose(w.compute_samples(t, sem), oracle.window(name, p, t, sem), rtol=0, atol=2e-16)


def test_tukey_rejects_degenerate_fractions():
    for f in (0.0, 1.0):
        with pytest.raises(fa.FNumberError) as e:
            fa.windows.tukey(f)
        assert e.value.identifier == "tukey:InvalidFractionsCosine"
