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


# ---- windowed normalization (windowed.m:81-86) -------------------------------
# Tapers whose value at t = NaN / inf is 0 (Tukey, boxcar) agree with MATLAB on
# element columns without positive rows (L = 0, where MATLAB forms 0/0); Hann and
# triangular tapers are checked with the Hann receive window, whose columns all
# have positive rows.
NW_CASES = [("tukey", ("tukey", 0.2)), ("boxcar", ("tukey", 0.2)), ("triangular", ("tukey", 0.2)),
            ("hann", ("tukey", 0.2)), ("tukey", ("boxcar", 0.0)), ("hann", ("hann", 0.0)),
            ("hann", ("triangular", 0.0))]


def _taper(fa, spec):
    name, p = spec
    return {"tukey": lambda: fa.windows.tukey(p), "boxcar": lambda: fa.windows.boxcar(),
            "hann": lambda: fa.windows.hann(), "triangular": lambda: fa.windows.triangular()}[name]()


@pytest.mark.parametrize("semantics", ["cpu", "mex"])
@pytest.mark.parametrize("win,taper", NW_CASES, ids=[f"{w}-{t[0]}{t[1]}" for w, t in NW_CASES])
def test_das_pw_windowed_normalization(fa, oracle, small_pw, win, taper, semantics):
    w, rf = small_pw
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    print(f"case das_pw windowed win={win} taper={taper} sem={semantics}", flush=True)
    img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width,
                    w.pitch, w.c_0, w.f_bounds, 0, _win(fa, win), Fo,
                    fa.normalizations.windowed(_taper(fa, taper)), fa.platforms.gpu(0), semantics=semantics,
                    output_scale="cpu")
    ref, info = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                              w.f_bounds, 0, (win, 0.5 if win == "tukey" else 0.0), _orc_F(oracle, Fo), "windowed",
                              semantics=semantics, norm_window=taper)
    assert np.all(np.isfinite(ref))
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL_L2 and d <= TOL_DB


@pytest.mark.parametrize("nw_tx,nw_rx", [(False, True), (True, False), (True, True)])
def test_das_sa_windowed_normalization(fa, oracle, nw_tx, nw_rx):
    from f_number_amd import synth

    w = synth.scaled(synth.saft_full(), 12, 10, n_samples=1024, n_speckle=600, n_elements=32)
    w.pitch = 4 * 0.3048e-3 / 1.0
    w.width = 4 * 0.27e-3
    rf = w.rf_host()
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    nwin = fa.windows.tukey(0.2)
    ntx = fa.normalizations.windowed(nwin) if nw_tx else fa.normalizations.on()
    nrx = fa.normalizations.windowed(nwin) if nw_rx else fa.normalizations.on()
    print(f"case das_sa windowed tx={nw_tx} rx={nw_rx}", flush=True)
    img = fa.das_sa(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.width, w.pitch, w.c_0,
                    w.f_bounds, 0, fa.windows.tukey(0.5), fa.windows.tukey(0.5), Fo, Fo, ntx, nrx,
                    output_scale="cpu")
    ref, info = oracle.das_sa(w.positions_x, w.positions_z, rf, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                              ("tukey", 0.5), ("tukey", 0.5), oracle.fnumber("angle_lb", 60, 1.5, 10),
                              oracle.fnumber("angle_lb", 60, 1.5, 10), "windowed" if nw_tx else "on",
                              "windowed" if nw_rx else "on", norm_window=("tukey", 0.2))
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL_L2 and d <= TOL_DB


def test_windowed_normalization_limits(fa, small_pw):
    """The taper path is FP32 and not differentiated: other precisions and the
    differentiation entry points report Unsupported."""
    w, rf = small_pw
    nrm = fa.normalizations.windowed(fa.windows.tukey(0.2))
    with pytest.raises(fa.FNumberError, match="Unsupported"):
        fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch, w.c_0,
                  w.f_bounds, 0.0, normalization=nrm, precision="fp16")
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, normalization=nrm)
    import torch

    x = torch.from_numpy(rf).cuda()
    img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    with pytest.raises(fa.FNumberError, match="Unsupported"):
        plan.execute_tangent(x, img, torch.empty_like(img))
    plan.close()


# ---- focused RF signal of the last voxel (das_pw.m:477-481, das_sa.m:501-505) --
@pytest.mark.parametrize("win,norm,sem", [("tukey", "on", "cpu"), ("boxcar", "off", "mex"), ("hann", "on", "cpu"),
                                          ("tukey", "windowed", "cpu")])
def test_focused_signal_matches_oracle(fa, oracle, small_pw, win, norm, sem):
    import torch

    w, rf = small_pw
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    nrm = {"on": fa.normalizations.on(), "off": fa.normalizations.off(),
           "windowed": fa.normalizations.windowed(fa.windows.tukey(0.2))}[norm]
    print(f"case focused signal win={win} norm={norm} sem={sem}", flush=True)
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=_win(fa, win), F_number=Fo, normalization=nrm, semantics=sem,
                     output_scale="cpu")
    x = torch.from_numpy(rf).cuda()
    img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    plan.execute(x, img)
    sig = plan.focused_signal().cpu().numpy()
    plan.close()
    _, info = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds,
                            0, (win, 0.5 if win == "tukey" else 0.0), _orc_F(oracle, Fo), norm, semantics=sem,
                            norm_window=("tukey", 0.2), want_signal=True)
    e = rel_l2(sig, info["signal"])
    print(f"  rel_l2={e:.3e} (N_dft {info['N_dft']})", flush=True)
    assert sig.shape == info["signal"].shape and e <= TOL_L2


def test_focused_signal_sa_matches_oracle(fa, oracle):
    import torch
    from f_number_amd import synth

    w = synth.scaled(synth.saft_full(), 12, 10, n_samples=1024, n_speckle=600, n_elements=32)
    w.pitch = 4 * 0.3048e-3 / 1.0
    w.width = 4 * 0.27e-3
    rf = w.rf_host()
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    print("case focused signal SA", flush=True)
    plan = fa.SaPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.f_s, w.width, w.pitch, w.c_0,
                     w.f_bounds, 0.0, window_tx=fa.windows.tukey(0.5), window_rx=fa.windows.tukey(0.5),
                     F_number_tx=Fo, F_number_rx=Fo, output_scale="cpu")
    x = torch.from_numpy(rf).cuda()
    img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    plan.execute(x, img)
    sig = plan.focused_signal().cpu().numpy()
    plan.close()
    _, info = oracle.das_sa(w.positions_x, w.positions_z, rf, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                            ("tukey", 0.5), ("tukey", 0.5), want_signal=True)
    e = rel_l2(sig, info["signal"])
    print(f"  rel_l2={e:.3e}", flush=True)
    assert e <= TOL_L2


def test_das_pw_nargout_matches_cpu_outputs(fa, oracle, small_pw):
    """[image, F_number_values, signal] = das_pw(...) (das_pw.m:1): F per bin and depth
    with the depth clamp, the signal of the last voxel, the image of the plain call."""
    w, rf = small_pw
    args = (w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, w.angles, w.width, w.pitch, w.c_0,
            w.f_bounds, 0, fa.windows.tukey(0.5), fa.f_numbers.grating.angle_lb(60, 1.5))
    img3, Fv, sig = fa.das_pw(*args, nargout=3, output_scale="cpu")
    img1 = fa.das_pw(*args, output_scale="cpu")
    img2, Fv2 = fa.das_pw(*args, nargout=2, output_scale="cpu")
    assert rel_l2(img3, img1) < 1e-6 and np.array_equal(Fv, Fv2)
    N = sig.size
    assert Fv.shape[1] == w.positions_z.size
    # the signal's first sample is the last voxel's image value / N_dft (sum of its focused spectrum)
    assert abs(N * sig[0] - img1[-1, -1]) <= 1e-5 * np.max(np.abs(img1))
    ref, info = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                              w.f_bounds, 0, ("tukey", 0.5), oracle.fnumber("angle_lb", 60, 1.5, 10), "on",
                              want_signal=True)
    assert rel_l2(sig, info["signal"]) <= TOL_L2


# ---- virtual-source transmits (diverging and focused waves) ------------------
VS = np.array([[-4e-3, -10e-3], [0.0, -15e-3], [5e-3, -8e-3], [0.0, 20e-3], [2e-3, 9e-3]])


@pytest.mark.parametrize("win,sem", [("tukey", "cpu"), ("hann", "cpu"), ("boxcar", "mex")])
def test_das_pw_virtual_sources_match_oracle(fa, oracle, small_pw, win, sem):
    w, _ = small_pw
    rng = np.random.default_rng(31)
    rf = rng.standard_normal((VS.shape[0], w.n_elements, w.n_samples)).astype(np.float32)
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    print(f"case virtual sources win={win} sem={sem}", flush=True)
    img = fa.das_pw(w.positions_x, w.positions_z, np.transpose(rf, (2, 1, 0)), w.f_s, np.zeros(VS.shape[0]), w.width,
                    w.pitch, w.c_0, w.f_bounds, 0, _win(fa, win), Fo, fa.normalizations.on(), fa.platforms.gpu(0),
                    semantics=sem, output_scale="cpu", virtual_sources=VS)
    ref, _ = oracle.das_pw(w.positions_x, w.positions_z, rf, w.f_s, np.zeros(VS.shape[0]), w.width, w.pitch, w.c_0,
                           w.f_bounds, 0, (win, 0.5 if win == "tukey" else 0.0), _orc_F(oracle, Fo), "on",
                           semantics=sem, virtual_sources=VS)
    e, d = rel_l2(img, ref), peak_db(img, ref)
    print(f"  rel_l2={e:.3e} peak={d:.1f} dB", flush=True)
    assert e <= TOL_L2 and d <= TOL_DB


def test_virtual_source_synth_focuses_points(fa):
    """Synthesizer (mode 2) and beamformer share the virtual-source model: point
    scatterers of a diverging-wave acquisition image at their positions."""
    import torch
    from f_number_amd import synth

    w = synth.scaled(synth.diverging_waves(), 96, 96, n_angles=5, n_speckle=0)
    w.n_points, w.snr_db = 6, np.inf
    xs, zs, _ = w.scatterers()
    print("case virtual-source synth", flush=True)
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=fa.windows.hann(), virtual_sources=w.virtual_sources)
    img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    plan.execute(w.rf_device("cuda:0"), img)
    a = img.abs().cpu().numpy()
    plan.close()
    dx, dz = np.diff(w.positions_x[:2])[0], np.diff(w.positions_z[:2])[0]
    for xp, zp in zip(xs, zs):
        ix, iz = int(round((xp - w.positions_x[0]) / dx)), int(round((zp - w.positions_z[0]) / dz))
        if not (3 <= ix < a.shape[0] - 3 and 3 <= iz < a.shape[1] - 3):
            continue
        win = a[ix - 3:ix + 4, iz - 3:iz + 4]
        jx, jz = np.unravel_index(np.argmax(win), win.shape)
        assert abs(jx - 3) <= 1 and abs(jz - 3) <= 1, (xp, zp)


@pytest.mark.parametrize("n_angles", [3, 40])
@pytest.mark.parametrize("win,F", [("boxcar", "constant"), ("tukey", "angle_lb")])
def test_short_bin_chunks_match_oracle(fa, oracle, win, F, n_angles):
    """Bin chunks of one, two and three bins (bin_split up to n_f): the receive apertures
    and support unions that lane groups compute once per four bins are clamped to the
    chunk's last bin.  Boxcar / F = 1.5 (support test per element) and Tukey / angle_lb
    (cut raised cosine) in FP32, one transmit tile (3 angles) and three (40 angles, lane
    swaps), each split against the oracle and against split 1; the tangent instance's
    image and dI/dc at each split against the forward image and split 1."""
    import torch

    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 20, 16, n_angles=n_angles, n_samples=1024, n_speckle=1500)
    if n_angles > 11:
        w.angles_deg = np.linspace(-10, 10, n_angles)
    rf_h = w.rf_host()
    Fo = fa.f_numbers.grating.angle_lb(60, 1.5) if F == "angle_lb" else fa.f_numbers.constant(1.5)
    ref, info = oracle.das_pw(w.positions_x, w.positions_z, rf_h, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                              w.f_bounds, 0, (win, 0.5 if win == "tukey" else 0.0), _orc_F(oracle, Fo), "on")
    n_f = info["n_ub"] - info["n_lb"] + 1
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=_win(fa, win), F_number=Fo, output_scale="cpu")
    rf = torch.from_numpy(rf_h).cuda()
    shape = (w.positions_x.size, w.positions_z.size)
    base = dbase = None
    for split in (1, 2, 3, (n_f + 2) // 3, (n_f + 1) // 2, n_f):
        plan.set_tiling(plan.info()["m_tiles"], split)
        img_d, t_d, dt_d = (torch.empty(shape, dtype=torch.complex64, device="cuda") for _ in range(3))
        plan.execute(rf, img_d)
        plan.execute_tangent(rf, t_d, dt_d)  # the tangent instance: image and dI/dc
        torch.cuda.synchronize()
        img, t_img, dimg = img_d.cpu().numpy().T, t_d.cpu().numpy().T, dt_d.cpu().numpy().T
        base = img if base is None else base
        dbase = dimg if dbase is None else dbase
        e, d, s = rel_l2(img, ref), peak_db(img, ref), rel_l2(img, base)
        et, sd = rel_l2(t_img, img), rel_l2(dimg, dbase)
        print(f"case {win}/{F} angles={n_angles} split={split} (n_f {n_f}): rel_l2={e:.3e} peak={d:.1f} dB "
              f"vs split 1 {s:.1e}; tangent image vs image {et:.1e}, dI/dc vs split 1 {sd:.1e}", flush=True)
        assert e <= TOL_L2 and d <= TOL_DB and s <= 1e-6
        assert et <= 1e-5 and sd <= 1e-5
    plan.close()


@pytest.mark.parametrize("form", ["sa_fp32", "pw_bf16x3", "pw_fp16"])
def test_short_bin_chunks_other_forms(fa, oracle, form):
    """The same bin chunks for the synthetic aperture's FP32 instance (receive and transmit
    apertures once per four bins, lane swaps) and the 16-bit plane-wave forms (lane
    permute): each split against the oracle (the form's own bar) and against split 1."""
    import torch

    from f_number_amd import synth

    Fo = fa.f_numbers.grating.angle_lb(60, 1.5)
    Fr = oracle.fnumber("angle_lb", 60, 1.5, 10)
    win = fa.windows.tukey(0.5)
    if form == "sa_fp32":
        w = synth.scaled(synth.saft_full(), 12, 10, n_samples=1024, n_speckle=600, n_elements=32)
        w.pitch = 4 * 0.3048e-3 / 1.0
        w.width = 4 * 0.27e-3
        rf_h = w.rf_host()
        ref, info = oracle.das_sa(w.positions_x, w.positions_z, rf_h, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds,
                                  0, ("tukey", 0.5), ("tukey", 0.5), Fr, Fr, "on", "on")
        plan = fa.SaPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.f_s, w.width, w.pitch, w.c_0,
                         w.f_bounds, 0.0, window_tx=win, window_rx=win, F_number_tx=Fo, F_number_rx=Fo,
                         output_scale="cpu")
        tol = TOL_L2
    else:
        prec = form.split("_")[1]
        w = synth.scaled(synth.cpwc_small(), 20, 16, n_angles=3, n_samples=1024, n_speckle=1500)
        rf_h = w.rf_host()
        ref, info = oracle.das_pw(w.positions_x, w.positions_z, rf_h, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                                  w.f_bounds, 0, ("tukey", 0.5), Fr, "on")
        plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width,
                         w.pitch, w.c_0, w.f_bounds, 0.0, window=win, F_number=Fo, precision=prec,
                         output_scale="cpu")
        tol = {"bf16x3": 1e-4, "fp16": 1e-3}[prec]
    n_f = info["n_ub"] - info["n_lb"] + 1
    rf = torch.from_numpy(rf_h).cuda()
    shape = (w.positions_x.size, w.positions_z.size)
    base = None
    for split in (1, 2, 3, (n_f + 2) // 3, (n_f + 1) // 2, n_f):
        plan.set_tiling(plan.info()["m_tiles"], split)
        img_d = torch.empty(shape, dtype=torch.complex64, device="cuda")
        plan.execute(rf, img_d)
        torch.cuda.synchronize()
        img = img_d.cpu().numpy().T
        base = img if base is None else base
        e, s = rel_l2(img, ref), rel_l2(img, base)
        print(f"case {form} split={split} (n_f {n_f}): rel_l2={e:.3e} vs split 1 {s:.1e}", flush=True)
        assert e <= tol and s <= 1e-6
    plan.close()


def test_split_images_are_reproducible_bitwise(fa):
    """Workgroups that share a voxel (8 bin chunks x 3 transmit groups here) write
    partial images that one kernel sums in a fixed order: two executions give the
    same bits, for the image and for the tangent dI/dc."""
    import torch

    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 30, 28, n_angles=11, n_samples=1024, n_speckle=1500)
    w.angles_deg = np.linspace(-10, 10, 40)
    rf = torch.from_numpy(w.rf_host()).cuda()
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=fa.windows.hann(), output_scale="cpu")
    plan.set_tiling(1, 8)
    info = plan.info()
    assert info["bin_split"] == 8 and info["m_tiles"] == 1
    shape = (w.positions_x.size, w.positions_z.size)
    a, b = (torch.empty(shape, dtype=torch.complex64, device="cuda") for _ in range(2))
    plan.execute(rf, a)
    plan.execute(rf, b)
    ta, tb, da, db = (torch.empty(shape, dtype=torch.complex64, device="cuda") for _ in range(4))
    plan.execute_tangent(rf, ta, da)
    plan.execute_tangent(rf, tb, db)
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(a, b) and torch.equal(ta, tb) and torch.equal(da, db)
    assert rel_l2(ta.cpu().numpy(), a.cpu().numpy()) < 1e-5
