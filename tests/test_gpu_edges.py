"""GPU parity on the reference's edge cases, against the FP64 oracle.

  empty apertures (das_pw.m:403,414; das_sa.m:417): voxels beside the array and just
      below it, where every bin's aperture is empty, give exactly zero (the bin is
      skipped; the reference MEX forms 0/0 there, das_kernel_normalization.cu:30-31),
      never NaN, in CPU and MEX semantics and every precision
  das_sa with distinct transmit and receive windows, F-numbers and normalizations,
      including transmit normalization off with receive on (gpu_bf_saft_rf.cu:24,
      1188-1199 give a zero image there; das_sa.m:479-480 normalize each side alone)
  a single depth (N_z = 1) and a single lateral position in MEX semantics (the N_pad
      of gpu_bf_das_pw_rf.cu:525 reads positions_z[1]; one depth has no second sample
      and falls back to the first)
Bars: FP32 rel-L2 <= 1e-4 and peak <= -80 dB; bf16x3 / fp16x3 <= 1e-4; fp16 <= 1e-3.
"""
import numpy as np
import pytest

from conftest import peak_db, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-4, "bf16x3": 1e-4, "fp16x3": 1e-4, "fp16": 1e-3}


@pytest.fixture(scope="module")
def fa():
    import f_number_amd as fa

    fa.load()
    return fa


def _win(fa, name):
    return {"boxcar": fa.windows.boxcar(), "hann": fa.windows.hann(), "tukey": fa.windows.tukey(0.5),
            "triangular": fa.windows.triangular()}[name]


def _wo(name):
    return (name, 0.5 if name == "tukey" else 0.0)


@pytest.fixture(scope="module")
def wide():
    """A grid reaching x = +-25 mm (the array spans +-19.35 mm) from z = 1 mm."""
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 26, 22, n_angles=3, n_samples=1024, n_speckle=1500)
    w.positions_x = np.linspace(-25e-3, 25e-3, 26)
    w.positions_z = np.linspace(1e-3, 30e-3, 22)
    return w, w.rf_host()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp16x3", "fp16"])
@pytest.mark.parametrize("semantics", ["cpu", "mex"])
@pytest.mark.parametrize("win,F", [("hann", "angle_lb"), ("tukey", "constant"), ("boxcar", "constant")])
def test_empty_apertures_are_exact_zeros(fa, oracle, wide, prec, semantics, win, F):
    w, rf = wide
    # a band above the F = 0 branch of angle_lb (p / lambda < 0.536, below 2.71 MHz, where
    # the aperture spans the whole array at every voxel): from 4 MHz, F >= 1.16, so the
    # voxels beside the array down to about 13 mm (x = +-25 mm), 8.5 mm (+-23 mm) and
    # 3.8 mm (+-21 mm) have no valid bin
    fb = (4.0e6, 6.75e6)
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5) if F == "angle_lb" else fa.f_numbers.constant(1.5)
    Fr = oracle.fnumber("angle_lb", 60, 1.5, 10) if F == "angle_lb" else oracle.fnumber("constant", 1.5)
    print(f"case empty apertures {win} {F} {semantics} {prec}", flush=True)
    img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width, w.pitch,
                    w.c_0, fb, 0, _win(fa, win), Fo, fa.normalizations.on(), fa.platforms.gpu(0),
                    semantics=semantics, precision=prec, output_scale="cpu")
    ref, _ = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, fb, 0,
                           _wo(win), Fr, "on", semantics=semantics)
    empty = ref == 0
    print(f"  {int(empty.sum())} of {empty.size} voxels have no valid bin", flush=True)
    assert empty.sum() >= 20
    assert np.all(np.isfinite(img))
    assert np.all(img[empty] == 0)
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL[prec]


SA_MIXES = [
    # (tx window, tx F, tx norm), (rx window, rx F, rx norm)
    (("hann", ("angle_lb",), "off"), ("tukey", ("constant", 1.0), "on")),
    (("boxcar", ("constant", 2.0), "off"), ("hann", ("angle_lb",), "on")),
    (("tukey", ("constant", 1.0), "on"), ("triangular", ("angle_lb",), "off")),
    (("triangular", ("constant", 0.0), "on"), ("boxcar", ("constant", 1.5), "on")),
]


def _sa_F(fa, oracle, f):
    if f[0] == "angle_lb":
        return fa.f_numbers.grating.angle_lb(60, 1.5), oracle.fnumber("angle_lb", 60, 1.5, 10)
    return fa.f_numbers.constant(f[1]), oracle.fnumber("constant", f[1])


@pytest.fixture(scope="module")
def sa_case():
    from f_number_amd import synth

    w = synth.scaled(synth.saft_full(), 14, 12, n_samples=1024, n_speckle=800, n_elements=32)
    w.pitch = 4 * 0.3048e-3
    w.width = 4 * 0.27e-3
    return w, w.rf_host()


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("semantics", ["cpu", "mex"])
@pytest.mark.parametrize("mix", range(len(SA_MIXES)))
def test_das_sa_distinct_tx_rx(fa, oracle, sa_case, mix, semantics, prec):
    w, rf = sa_case
    tx, rx = SA_MIXES[mix]
    Ftx, Ftx_o = _sa_F(fa, oracle, tx[1])
    Frx, Frx_o = _sa_F(fa, oracle, rx[1])
    nrm = {"on": fa.normalizations.on(), "off": fa.normalizations.off()}
    print(f"case das_sa tx={tx} rx={rx} {semantics} {prec}", flush=True)
    img = fa.das_sa(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.width, w.pitch, w.c_0,
                    w.f_bounds, 0, _win(fa, tx[0]), _win(fa, rx[0]), Ftx, Frx, nrm[tx[2]], nrm[rx[2]],
                    semantics=semantics, precision=prec, output_scale="cpu")
    ref, _ = oracle.das_sa(w.positions_x, w.positions_z, rf, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                           _wo(tx[0]), _wo(rx[0]), Ftx_o, Frx_o, tx[2], rx[2], semantics=semantics)
    assert np.linalg.norm(ref) > 0
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert np.all(np.isfinite(img))
    assert e <= TOL[prec]
    if prec == "fp32":
        assert d <= -80.0


@pytest.mark.parametrize("semantics", ["mex", "cpu"])
@pytest.mark.parametrize("shape", ["one_depth", "one_line"])
def test_single_depth_or_line(fa, oracle, semantics, shape):
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 40, 40, n_angles=5, n_samples=1024, n_speckle=1500)
    rf = w.rf_host()
    px, pz = w.positions_x, w.positions_z
    if shape == "one_depth":
        pz = np.array([17.3e-3])
    else:
        px = np.array([1.1e-3])
    print(f"case {shape} {semantics}: {pz.size} x {px.size}", flush=True)
    img = fa.das_pw(px, pz, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                    fa.windows.tukey(0.5), fa.f_numbers.grating.angle_lb(60, 1.5), fa.normalizations.on(),
                    fa.platforms.gpu(0), semantics=semantics, output_scale="cpu")
    ref, info = oracle.das_pw(px, pz, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0, ("tukey", 0.5),
                              oracle.fnumber("angle_lb", 60, 1.5, 10), "on", semantics=semantics)
    assert img.shape == ref.shape == (pz.size, px.size)
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  N_dft={info['N_dft']} rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= 1e-4 and d <= -80.0
