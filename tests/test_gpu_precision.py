"""GPU parity of the mixed-precision contractions against the FP64 oracle.

The receive apodization x phase tile and the spectra enter the matrix cores as
16-bit operands with FP32 accumulation; delays and phase reduction stay FP64
(integer phase wrap), windows and normalisation FP32.

Bars (relative L2 error of the image vs the FP64 restatement of the CPU path,
same inputs):
  bf16x3, fp16x3 (hi + lo split, three passes)  <= 1e-4  (the FP32 bar)
  fp16 (one pass, per-bin power-of-two scaling)  <= 1e-3
Each case prints its measured error.
"""
import numpy as np
import pytest

from conftest import peak_db, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"bf16x3": 1e-4, "fp16x3": 1e-4, "fp16": 1e-3}
PRECS = ["bf16x3", "fp16x3", "fp16"]


@pytest.fixture(scope="module")
def fa():
    import f_number_amd as fa

    fa.load()
    return fa


def _win(fa, name):
    return {"boxcar": fa.windows.boxcar(), "hann": fa.windows.hann(), "tukey": fa.windows.tukey(0.5),
            "triangular": fa.windows.triangular()}[name]


def _pw_case(n_el, n_ang, seed):
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_full(), 14, 12, n_angles=n_ang, n_samples=1024, n_speckle=1500, n_elements=n_el)
    w.seed = seed
    return w, w.rf_host()


_PW_SHAPES = {}


def _pw(n_el, n_ang):
    key = (n_el, n_ang)
    if key not in _PW_SHAPES:
        _PW_SHAPES[key] = _pw_case(n_el, n_ang, 9100 + n_el + n_ang)
    return _PW_SHAPES[key]


def _run_pw(fa, oracle, w, rf, win, prec, semantics="cpu", norm="on"):
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    nrm = fa.normalizations.on() if norm == "on" else fa.normalizations.off()
    img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width,
                    w.pitch, w.c_0, w.f_bounds, 0, _win(fa, win), Fo, nrm, fa.platforms.gpu(0),
                    semantics=semantics, precision=prec, output_scale="cpu")
    ref, _ = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                           w.f_bounds, 0, (win, 0.5 if win == "tukey" else 0.0),
                           oracle.fnumber("angle_lb", 60, 1.5, 10), norm, semantics=semantics)
    return img, ref


# element counts select the 8/16/32-element-per-lane variants; angle counts
# cover 1, 2, 3 and 5 transmit tiles of 16 rows (75 angles: one group of 5 tiles)
@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_el,n_ang", [(128, 3), (64, 20), (32, 40), (100, 75)])
def test_das_pw_precision_shapes(fa, oracle, prec, n_el, n_ang):
    w, rf = _pw(n_el, n_ang)
    print(f"case das_pw prec={prec} n_el={n_el} n_ang={n_ang}", flush=True)
    img, ref = _run_pw(fa, oracle, w, rf, "hann", prec)
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("win,semantics,norm", [("boxcar", "cpu", "on"), ("tukey", "mex", "on"),
                                                ("triangular", "cpu", "off"), ("hann", "mex", "off")])
def test_das_pw_precision_windows(fa, oracle, prec, win, semantics, norm):
    w, rf = _pw(128, 3)
    print(f"case das_pw prec={prec} win={win} sem={semantics} norm={norm}", flush=True)
    img, ref = _run_pw(fa, oracle, w, rf, win, prec, semantics, norm)
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL[prec]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n_el,win", [(32, "hann"), (32, "tukey"), (64, "boxcar")])
def test_das_sa_precision(fa, oracle, prec, n_el, win):
    from f_number_amd import synth

    w = synth.scaled(synth.saft_full(), 12, 10, n_samples=1024, n_speckle=600, n_elements=n_el)
    w.pitch = 4 * 0.3048e-3
    w.width = 4 * 0.27e-3
    rf = w.rf_host()
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    print(f"case das_sa prec={prec} n_el={n_el} win={win}", flush=True)
    img = fa.das_sa(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.width, w.pitch, w.c_0,
                    w.f_bounds, 0, _win(fa, win), _win(fa, win), Fo, Fo, fa.normalizations.on(),
                    fa.normalizations.on(), precision=prec, output_scale="cpu")
    wt = (win, 0.5 if win == "tukey" else 0.0)
    Fr = oracle.fnumber("angle_lb", 60, 1.5, 10)
    ref, _ = oracle.das_sa(w.positions_x, w.positions_z, rf, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                           wt, wt, Fr, Fr, "on", "on")
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL[prec]


def test_precision_zero_input(fa):
    """All-zero RF: the per-bin scale falls back to 1 and the image is exactly zero."""
    w, rf = _pw(64, 20)
    z = np.zeros_like(rf)
    for prec in PRECS:
        img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(z, (2, 1, 0)), w.f_s, w.angles, w.width,
                        w.pitch, w.c_0, w.f_bounds, precision=prec)
        assert np.all(img == 0)


def test_precision_scale_invariance(fa):
    """fp16 with per-bin power-of-two scaling: scaling the RF by 2^20 scales the image by 2^20."""
    w, rf = _pw(64, 20)
    args = (w.positions_x, w.positions_z)
    a = fa.das_pw(*args, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds,
                  precision="fp16")
    b = fa.das_pw(*args, np.transpose(rf * 2.0**20, (2, 1, 0)), w.f_s, w.angles, w.width, w.pitch, w.c_0,
                  w.f_bounds, precision="fp16")
    assert rel_l2(b / 2.0**20, a) < 1e-6
