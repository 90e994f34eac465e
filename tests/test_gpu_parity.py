"""GPU parity: libfnumber_hip (HIP, gfx950) against the FP64 oracle.

Bar (plain FP32 path vs the reference CPU path restated in FP64): relative L2
error <= 1e-4 and peak-normalised max error <= -80 dB, after bringing the GPU
image to the CPU path's scale (output_scale='cpu': 2 * sum, das_pw.m:384,467).
"""
import numpy as np
import pytest

from conftest import peak_db, rel_l2

pytestmark = pytest.mark.gpu

TOL_L2 = 1e-4
TOL_DB = -80.0


@pytest.fixture(scope="module")
def fa():
    import f_number_amd as fa

    fa.load()
    return fa


@pytest.fixture(scope="module")
def small_pw():
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 24, 20, n_angles=3, n_samples=1024, n_speckle=2000)
    return w, w.rf_host()


def _win(fa, name):
    return {"boxcar": fa.windows.boxcar(), "hann": fa.windows.hann(), "tukey": fa.windows.tukey(0.5),
            "triangular": fa.windows.triangular()}[name]


def _orc_F(oracle, F):
    import f_number_amd as fa

    if isinstance(F, fa.f_numbers.grating.angle_lb):
        return oracle.fnumber("angle_lb", F.angle_lb_deg, F.F_number_ub, F.distance_deg)
    return oracle.fnumber("constant", F.value)


@pytest.mark.parametrize("semantics", ["cpu", "mex"])
@pytest.mark.parametrize("norm", ["on", "off"])
@pytest.mark.parametrize("F", ["angle_lb", "constant"])
@pytest.mark.parametrize("win", ["tukey", "boxcar", "hann", "triangular"])
def test_das_pw_matches_oracle(fa, oracle, small_pw, win, F, norm, semantics):
    w, rf = small_pw
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5) if F == "angle_lb" else fa.f_numbers.constant(1.5)
    nrm = fa.normalizations.on() if norm == "on" else fa.normalizations.off()
    print(f"case das_pw win={win} F={F} norm={norm} sem={semantics}", flush=True)
    img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width,
                    w.pitch, w.c_0, w.f_bounds, 0, _win(fa, win), Fo, nrm, fa.platforms.gpu(0),
                    semantics=semantics, output_scale="cpu")
    ref, info = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                              w.f_bounds, 0, (win, 0.5 if win == "tukey" else 0.0), _orc_F(oracle, Fo), norm,
                              semantics=semantics)
    assert img.shape == ref.shape
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB N_dft={info['N_dft']}", flush=True)
    assert e <= TOL_L2 and d <= TOL_DB


def test_das_pw_mex_scale(fa, oracle, small_pw):
    """MEX output = CPU output / N_dft (gpu_bf_das_pw_rf.cu:1118 vs das_pw.m:384)."""
    w, rf = small_pw
    args = (w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width, w.pitch, w.c_0,
            w.f_bounds)
    a = fa.das_pw(*args, output_scale="mex")
    b, wts, info = fa.das_pw(*args, output_scale="cpu", return_weights=True)
    assert rel_l2(a * info["N_dft"], b) < 1e-6


def test_das_pw_single_angle_2d(fa, oracle, small_pw):
    """One angle with 2-D RF is the reference call exactly; compounding = sum of single calls."""
    w, rf = small_pw
    tot = 0
    for a in range(w.angles.size):
        tot = tot + fa.das_pw(w.positions_x, w.positions_z, rf[a].T, w.f_s, float(w.angles[a]), w.width,
                              w.pitch, w.c_0, w.f_bounds, output_scale="cpu")
    comp = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width,
                     w.pitch, w.c_0, w.f_bounds, output_scale="cpu")
    assert rel_l2(comp, tot) < 1e-5


@pytest.mark.parametrize("norm", ["on", "off"])
@pytest.mark.parametrize("win", ["tukey", "boxcar", "hann"])
def test_das_sa_matches_oracle(fa, oracle, win, norm):
    from f_number_amd import synth

    w = synth.scaled(synth.saft_full(), 12, 10, n_samples=1024, n_speckle=600, n_elements=32)
    w.pitch = 4 * 0.3048e-3 / 1.0
    w.width = 4 * 0.27e-3
    rf = w.rf_host()
    nrm = fa.normalizations.on() if norm == "on" else fa.normalizations.off()
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    print(f"case das_sa win={win} norm={norm}", flush=True)
    img = fa.das_sa(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.width, w.pitch, w.c_0,
                    w.f_bounds, 0, _win(fa, win), _win(fa, win), Fo, Fo, nrm, nrm, output_scale="cpu")
    wt = (win, 0.5 if win == "tukey" else 0.0)
    ref, info = oracle.das_sa(w.positions_x, w.positions_z, rf, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                              wt, wt, oracle.fnumber("angle_lb", 60, 1.5, 10), oracle.fnumber("angle_lb", 60, 1.5, 10),
                              norm, norm)
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL_L2 and d <= TOL_DB


def test_das_pw_full_size_voxel_subset(fa, oracle):
    """11-angle 256 x 256 CPWC at full size, checked on a seeded voxel subset."""
    from f_number_amd import synth

    w = synth.cpwc_small()
    w.n_speckle = 4000
    rf = w.rf_host()
    win = fa.windows.tukey(0.5)
    img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width,
                    w.pitch, w.c_0, w.f_bounds, 0, win, output_scale="cpu")
    rng = np.random.default_rng(7)
    ix = rng.integers(0, w.positions_x.size, 96)
    iz = rng.integers(0, w.positions_z.size, 96)
    ref, info = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                              w.f_bounds, 0, ("tukey", 0.5), oracle.fnumber("angle_lb", 60, 1.5, 10), "on",
                              sel=(ix, iz))
    got = img[iz, ix]
    e = rel_l2(got, ref)
    print(f"  N_dft={info['N_dft']} band={info['n_lb']}..{info['n_ub']} rel_l2={e:.3e}", flush=True)
    assert info["N_dft"] == 2509 and info["n_lb"] == 271 and info["n_ub"] == 812
    assert e <= TOL_L2
