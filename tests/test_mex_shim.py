"""The MEX gateways (f_number_amd/csrc/mex/) driven through a fake MEX runtime:
the reference's positional argument lists, defaults, error identifiers and
output formats (gpu_bf_das_pw_rf.cu:110-485, gpu_bf_saft_rf.cu:107-592)."""
import ctypes as C

import numpy as np
import pytest

import mexsim
from conftest import rel_l2


def _devices():
    import f_number_amd as fa

    n = C.c_int32()
    fa.load().fn_device_count(C.byref(n))
    return n.value


@pytest.fixture(scope="module")
def pw():
    mexsim.build()
    return mexsim.load("pw")


@pytest.fixture(scope="module")
def sa():
    mexsim.build()
    return mexsim.load("sa")


def _pw_args(g, rf, w, angle):
    return [g.double(w.positions_x), g.double(w.positions_z), g.double(rf), g.double(w.f_s), g.double(angle),
            g.double(w.width), g.double(w.pitch), g.double(w.c_0)]


def test_gateway_reports_missing_device_first(pw):
    """With no GPU the gateway stops at its first check (gpu_bf_das_pw_rf.cu:115-121)."""
    if _devices() > 0:
        pytest.skip("a GPU is present")
    out, err = pw.call(1, [pw.empty()] * 8)
    assert out is None and err == "gpu_bf_das_pw_rf:NoCUDADevices"


@pytest.mark.gpu
def test_pw_gateway_argument_errors(pw):
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 4, 4, n_angles=1, n_samples=256, n_speckle=0)
    rf = np.zeros((256, 16))
    base = _pw_args(pw, rf, w, 0.0)
    assert pw.call(1, base[:7])[1] == "gpu_bf_das_pw_rf:nrhs"
    assert pw.call(3, base)[1] == "gpu_bf_das_pw_rf:nlhs"
    bad = list(base)
    bad[2] = pw.double(np.zeros((4, 4, 4, 4)))
    assert pw.call(1, bad)[1] == "gpu_bf_das_pw_rf:InvalidRFData"
    bad = list(base)
    bad[4] = pw.double(2.0)
    assert pw.call(1, bad)[1] == "gpu_bf_das_pw_rf:InvalidSteeringAngle"
    bad = list(base)
    bad[6] = pw.double(0.1e-3)
    assert pw.call(1, bad)[1] == "gpu_bf_das_pw_rf:InvalidElementPitch"
    assert pw.call(1, base + [pw.double([6e6, 5e6])])[1] == "gpu_bf_das_pw_rf:InvalidFrequencyBound"
    assert pw.call(1, base + [pw.empty(), pw.double(-1.0)])[1] == "gpu_bf_das_pw_rf:InvalidTimeShift"
    assert pw.call(1, base + [pw.empty(), pw.empty(), pw.double(3.0)])[1] == "gpu_bf_das_pw_rf:InvalidWindow"
    assert pw.call(1, base + [pw.empty()] * 3 + [pw.window("hann")])[1] == "gpu_bf_das_pw_rf:InvalidFNumber"
    assert pw.call(1, base + [pw.empty()] * 5 + [pw.double(99.0)])[1] == "gpu_bf_das_pw_rf:InvalidCUDADevice"
    assert pw.call(1, base + [pw.empty()] * 6 + [pw.double(7.0)])[1] == "gpu_bf_das_pw_rf:InvalidVerbosityLevel"


@pytest.mark.gpu
def test_pw_gateway_matches_python_api(pw):
    import f_number_amd as fa
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 20, 16, n_angles=3, n_samples=1024, n_speckle=800)
    rf = w.rf_host()
    rf_m = np.transpose(rf, (2, 1, 0)).astype(np.float64)  # MATLAB [N_t x N_el x N_a]
    args = [pw.double(w.positions_x), pw.double(w.positions_z), pw.double(rf_m), pw.double(w.f_s),
            pw.double(w.angles), pw.double(w.width), pw.double(w.pitch), pw.double(w.c_0),
            pw.double(list(w.f_bounds)), pw.double(0.0), pw.window("tukey", fraction_cosine=0.5),
            pw.fnumber("f_numbers.grating.angle_lb", angle_lb_deg=60.0, F_number_ub=1.5, distance_deg=10.0),
            pw.normalization("on"), pw.double(0.0), pw.double(0.0)]
    out, err = pw.call(2, args)
    assert err is None
    img, wts = out
    ref, wref, info = fa.das_pw(w.positions_x, w.positions_z, rf_m, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                                w.f_bounds, 0, fa.windows.tukey(0.5), fa.f_numbers.grating.angle_lb(60, 1.5),
                                return_weights=True)
    assert img.shape == (w.positions_z.size, w.positions_x.size) and img.dtype == np.complex64
    assert rel_l2(img, ref) < 1e-6 and np.allclose(wts, wref)
    # defaults of the gateway: boxcar, angle_lb(70, 1.5, 10), normalization on
    out, err = pw.call(1, args[:8] + [pw.double(list(w.f_bounds))])
    ref = fa.das_pw(w.positions_x, w.positions_z, rf_m, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                    fa.windows.boxcar(), fa.f_numbers.grating.angle_lb(70, 1.5, 10))
    assert err is None and rel_l2(out[0], ref) < 1e-6
    # 'mex' semantics selects the reference kernels' choices
    out_m, err = pw.call(1, args[:13] + [pw.double(0.0), pw.double(0.0), pw.string("mex")])
    ref_m = fa.das_pw(w.positions_x, w.positions_z, rf_m, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                      fa.windows.tukey(0.5), fa.f_numbers.grating.angle_lb(60, 1.5), semantics="mex")
    assert err is None and rel_l2(out_m[0], ref_m) < 1e-6


@pytest.mark.gpu
def test_sa_gateway_matches_python_api(sa):
    import f_number_amd as fa
    from f_number_amd import synth

    w = synth.scaled(synth.saft_full(), 10, 8, n_samples=768, n_speckle=300, n_elements=16)
    rf = w.rf_host()
    rf_m = np.transpose(rf, (2, 1, 0)).astype(np.float64)
    F = sa.fnumber("f_numbers.grating.angle_lb", angle_lb_deg=60.0, F_number_ub=1.5, distance_deg=10.0)
    args = [sa.double(w.positions_x), sa.double(w.positions_z), sa.double(rf_m), sa.double(w.f_s),
            sa.double(w.width), sa.double(w.pitch), sa.double(w.c_0), sa.double(list(w.f_bounds)), sa.double(0.0),
            sa.window("hann"), sa.window("hann"), F, F, sa.normalization("on"), sa.normalization("on")]
    out, err = sa.call(3, args)
    assert err is None
    ref, wt, wr, info = fa.das_sa(w.positions_x, w.positions_z, rf_m, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                                  fa.windows.hann(), fa.windows.hann(), fa.f_numbers.grating.angle_lb(60, 1.5),
                                  fa.f_numbers.grating.angle_lb(60, 1.5), return_weights=True)
    assert rel_l2(out[0], ref) < 1e-6 and np.allclose(out[1], wt) and np.allclose(out[2], wr)
    bad = list(args)
    bad[2] = sa.double(np.zeros((64, 4, 5)))
    assert sa.call(1, bad)[1] == "gpu_bf_saft_rf:InvalidRFData"


@pytest.mark.gpu
def test_gateways_report_at_verbosity_one(pw, sa, capfd):
    """Verbosity 1 prints the status section before the computation (geometry, band
    with its indices, names of window, F-number and normalization through MATLAB's
    string and char, gpu_bf_das_pw_rf.cu:609-632) and, after it, the device, launch,
    memory and timing sections (gpu_bf_das_pw_rf.cu:664-689, 875-903, 1128-1144);
    verbosity 0 prints nothing."""
    import re

    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 12, 10, n_angles=3, n_samples=1024, n_speckle=300)
    rf_m = np.transpose(w.rf_host(), (2, 1, 0)).astype(np.float64)
    args = [pw.double(w.positions_x), pw.double(w.positions_z), pw.double(rf_m), pw.double(w.f_s),
            pw.double(w.angles), pw.double(w.width), pw.double(w.pitch), pw.double(w.c_0),
            pw.double(list(w.f_bounds)), pw.double(0.0), pw.window("tukey", fraction_cosine=0.5),
            pw.fnumber("f_numbers.grating.angle_lb", angle_lb_deg=60.0, F_number_ub=1.5, distance_deg=10.0),
            pw.normalization("on"), pw.double(0.0)]
    capfd.readouterr()
    assert pw.call(1, args + [pw.double(0.0)])[1] is None
    assert capfd.readouterr().out == ""
    assert pw.call(1, args + [pw.double(1.0)])[1] is None
    out = capfd.readouterr().out
    print(out)
    for section in ("status information", "lattice geometry", "lower frequency bound", "Selected device 0 of",
                    "launch parameters", "memory consumption", "performance statistics"):
        assert section in out, section
    assert "windows.tukey (fraction_cosine = 0.50)" in out
    assert "f_numbers.grating.angle_lb (" in out and "angle_lb_deg = 60.00" in out
    assert re.search(r"lattice geometry\s*:\s*12 x 10", out)
    assert out.index("status information") < out.index("Selected device")
    times = {k: float(v) for k, v in re.findall(r"^\s+(memory transfer: host to device|band-limited DFT \(spectra\)|"
                                                 r"beamforming kernel|memory transfer: device to host|total)\s*:\s*"
                                                 r"([0-9.]+) ms", out, re.M)}
    assert len(times) == 5 and times["beamforming kernel"] > 0
    assert abs(sum(v for k, v in times.items() if k != "total") - times["total"]) < 2e-3 + 1e-3 * times["total"]
    mem = out[out.index("memory consumption"):out.index("performance statistics")]
    kib = re.findall(r"^\s+(.+?)\s*:\s+([0-9.]+) KiByte", mem, re.M)
    parts = [float(v) for k, v in kib if k != "total"]
    total = float(dict(kib)["total"])
    assert abs(sum(parts) - total) < 0.05 * len(parts)
    rf_kib = dict(kib)["RF signals"]
    assert abs(float(rf_kib) - rf_m.size * 8 / 1024) < 0.01
    # the synthetic-aperture gateway prints both sides' names
    ws = synth.scaled(synth.saft_full(), 6, 5, n_samples=512, n_speckle=100, n_elements=16)
    rs = np.transpose(ws.rf_host(), (2, 1, 0)).astype(np.float64)
    F = sa.fnumber("f_numbers.grating.angle_lb", angle_lb_deg=60.0, F_number_ub=1.5, distance_deg=10.0)
    sargs = [sa.double(ws.positions_x), sa.double(ws.positions_z), sa.double(rs), sa.double(ws.f_s),
             sa.double(ws.width), sa.double(ws.pitch), sa.double(ws.c_0), sa.double(list(ws.f_bounds)),
             sa.double(0.0), sa.window("hann"), sa.window("boxcar"), F, F, sa.normalization("on"),
             sa.normalization("off"), sa.double(0.0), sa.double(1.0)]
    assert sa.call(1, sargs)[1] is None
    out = capfd.readouterr().out
    for section in ("window (tx)", "windows.hann", "window (rx)", "windows.boxcar", "normalization (rx)",
                    "normalizations.off", "Selected device 0 of", "performance statistics"):
        assert section in out, section


@pytest.fixture(scope="module")
def pwadj():
    mexsim.build()
    return mexsim.load("pwadj")


def test_adjoint_gateway_reports_missing_device_first(pwadj):
    if _devices() > 0:
        pytest.skip("a GPU is present")
    out, err = pwadj.call(1, [pwadj.empty()] * 9)
    assert out is None and err == "gpu_bf_das_pw_rf_adjoint:NoCUDADevices"


@pytest.mark.gpu
def test_adjoint_gateway_is_the_adjoint_of_the_forward_gateway(pw, pwadj):
    """<A x, y> = <x, A^H y> across the two MEX gateways (MEX output scale, single
    precision outputs, sums in FP64), with delay offsets and c_tof through the adjoint
    gateway and the Python plan API for the matching forward."""
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 12, 10, n_angles=3, n_samples=768, n_speckle=600, n_elements=32)
    rf = w.rf_host()
    rf_m = np.transpose(rf, (2, 1, 0)).astype(np.float64)
    common = [pw.double(w.f_s), pw.double(w.angles), pw.double(w.width), pw.double(w.pitch), pw.double(w.c_0),
              pw.double(list(w.f_bounds)), pw.double(0.0), pw.window("hann"),
              pw.fnumber("f_numbers.grating.angle_lb", angle_lb_deg=60.0, F_number_ub=1.5, distance_deg=10.0),
              pw.normalization("on"), pw.double(0.0), pw.double(0.0)]
    out, err = pw.call(1, [pw.double(w.positions_x), pw.double(w.positions_z), pw.double(rf_m)] + common)
    assert err is None
    Ax = out[0].astype(np.complex128)
    rng = np.random.default_rng(2)
    y = (Ax + 0.1 * np.abs(Ax).mean() * (rng.normal(size=Ax.shape) + 1j * rng.normal(size=Ax.shape)))
    y = y.astype(np.complex64)
    g = pwadj
    cargs = [g.double(w.f_s), g.double(w.angles), g.double(w.width), g.double(w.pitch), g.double(w.c_0),
             g.double(list(w.f_bounds)), g.double(0.0), g.window("hann"),
             g.fnumber("f_numbers.grating.angle_lb", angle_lb_deg=60.0, F_number_ub=1.5, distance_deg=10.0),
             g.normalization("on"), g.double(0.0), g.double(0.0)]
    out, err = g.call(1, [g.double(w.positions_x), g.double(w.positions_z), g.single_complex(y),
                          g.double([w.n_samples, w.n_elements])] + cargs)
    assert err is None
    AHy = out[0]
    assert AHy.shape == (w.n_samples, w.n_elements, 3)
    lhs = np.real(np.vdot(Ax, y.astype(np.complex128)))
    rhs = np.sum(rf_m * AHy)
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    # with delay offsets and c_tof: matches the plan API's adjoint (MEX scale)
    delays = rng.normal(0, 20e-9, w.n_elements)
    out, err = g.call(1, [g.double(w.positions_x), g.double(w.positions_z), g.single_complex(y),
                          g.double([w.n_samples, w.n_elements])] + cargs +
                      [g.double(1510.0), g.double(delays)])
    assert err is None
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=fa.windows.hann(), F_number=fa.f_numbers.grating.angle_lb(60, 1.5),
                     normalization=fa.normalizations.on(), output_scale="mex")
    plan.set_tof(1510.0, delays)
    yd = torch.from_numpy(np.ascontiguousarray(y.T)).cuda()
    ref = plan.adjoint(yd)
    torch.cuda.synchronize()
    ref = np.transpose(ref.cpu().numpy(), (2, 1, 0))
    plan.close()
    assert rel_l2(out[0], ref) < 1e-6
