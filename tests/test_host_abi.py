"""The C ABI of libfnumber_hip.so without a GPU: exported symbols, host-side
configuration logic and the argument checks with the reference's error ids."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import f_number_amd as fa
from f_number_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "fnumber_hip.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(fn_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
        assert n in _lib.PROTOTYPES, n
    assert L.fn_abi_version() == 1


def test_status_ids_follow_the_mex():
    L = _lib.load()
    assert L.fn_status_id(6).decode() == "InvalidRFData"
    assert L.fn_status_id(12).decode() == "InvalidFrequencyBounds"
    assert L.fn_status_id(19).decode() == "InvalidCUDADevice"


def _params(**kw):
    from f_number_amd.das import pw_params

    base = dict(px=np.linspace(-1e-3, 1e-3, 4), pz=np.linspace(5e-3, 8e-3, 5), n_t=256, n_el=16, angles=[0.0],
                f_s=20.832e6, w=0.27e-3, p=0.3048e-3, c=1540.0, fb=(2e6, 7e6), t0=0.0, window=fa.windows.hann(),
                F=fa.f_numbers.constant(1.5), normalization=fa.normalizations.on())
    base.update(kw)
    return pw_params(**base)


@pytest.mark.parametrize("field,value,ident", [
    ("f_s", -1.0, "InvalidSamplingRate"),
    ("element_width", 0.0, "InvalidElementWidth"),
    ("element_pitch", 0.1e-3, "InvalidElementPitch"),
    ("c_0", float("nan"), "InvalidSoundSpeed"),
    ("index_t0", -1.0, "InvalidTimeShift"),
    ("verbosity", 4, "InvalidVerbosityLevel"),
    ("index_gpu", -1, "InvalidCUDADevice"),
])
def test_validation_error_ids(field, value, ident):
    P, keep, info = _params()
    setattr(P, field, value)
    st = _lib.load().fn_pw_validate(C.byref(P))
    assert _lib.load().fn_status_id(st).decode() == ident


def test_validation_f_values_and_windows():
    L = _lib.load()
    P, keep, info = _params()
    assert L.fn_pw_validate(C.byref(P)) == 0
    F = np.array(keep[3])
    F[2] = -1.0
    P.f_number_values = F.ctypes.data_as(C.POINTER(C.c_double))
    assert L.fn_status_id(L.fn_pw_validate(C.byref(P))).decode() == "InvalidFNumber"
    P, keep, info = _params()
    P.n_f = P.n_f - 1
    assert L.fn_status_id(L.fn_pw_validate(C.byref(P))).decode() == "InvalidFNumber"
    P, keep, info = _params()
    P.window = _lib.Window(7, 0.0)
    assert L.fn_status_id(L.fn_pw_validate(C.byref(P))).decode() == "InvalidWindow"
    P, keep, info = _params()
    P.window = _lib.Window(2, 1.5)
    assert L.fn_status_id(L.fn_pw_validate(C.byref(P))).decode() == "InvalidWindow"
    P, keep, info = _params()
    ang = np.array([2.0])
    P.steering_angles = ang.ctypes.data_as(C.POINTER(C.c_double))
    assert L.fn_status_id(L.fn_pw_validate(C.byref(P))).decode() == "InvalidSteeringAngle"


def test_band_checks_follow_the_mex():
    with pytest.raises(fa.FNumberError) as e:
        _params(fb=(9.0e6, 9.5e6), f_s=18e6)
    assert e.value.identifier == "gpu_bf_das_pw_rf:InvalidFrequencyBound"


@pytest.mark.parametrize("kw,ident", [
    (dict(positions_x=np.array([[0.0, 1.0], [1.0, 2.0]])), "das_pw:NoNumericAndRealRowVector"),
    (dict(positions_x=np.array([0.0, np.inf])), "das_pw:NoFiniteXCoordinates"),
    (dict(positions_z=np.array([0.0, 1e-3])), "das_pw:NoPositiveAndFiniteZCoordinates"),
    (dict(data_RF=np.zeros((1, 4))), "das_pw:NoSamples"),
    (dict(data_RF=np.full((8, 4), np.nan)), "das_pw:NoFiniteRFData"),
    (dict(f_s=0.0), "das_pw:InvalidSamplingRate"),
    (dict(steering_angle=1.6), "das_pw:InvalidSteeringAngle"),
    (dict(element_pitch=0.2e-3), "das_pw:InvalidElementPitch"),
    (dict(c_0=-1.0), "das_pw:InvalidSpeedOfSound"),
    (dict(f_bounds=[5e6, 4e6]), "das_pw:InvalidFrequencyBounds"),
    (dict(index_t0=1.5), "das_pw:InvalidTimeIndex"),
    (dict(window="hann"), "das_pw:NoWindow"),
    (dict(F_number=1.5), "das_pw:NoFNumber"),
    (dict(normalization="on"), "das_pw:NoNormalization"),
    (dict(platform=fa.platforms.cpu()), "das_pw:UnsupportedPlatform"),
])
def test_das_pw_argument_checks(kw, ident):
    args = dict(positions_x=np.linspace(-1e-3, 1e-3, 4), positions_z=np.linspace(5e-3, 8e-3, 5),
                data_RF=np.zeros((64, 8)), f_s=20.832e6, steering_angle=0.0, element_width=0.27e-3,
                element_pitch=0.3048e-3, c_0=1540.0)
    args.update(kw)
    with pytest.raises(fa.FNumberError) as e:
        fa.das_pw(**args)
    assert e.value.identifier == ident


def test_das_sa_argument_checks():
    with pytest.raises(fa.FNumberError) as e:
        fa.das_sa(np.zeros(3), np.ones(3) * 1e-3, np.zeros((64, 8, 7)), 20e6, 0.27e-3, 0.3e-3, 1540.0)
    assert e.value.identifier == "gpu_bf_saft_rf:InvalidRFData"


def test_synth_host_is_deterministic_and_focused():
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 4, 4, n_angles=2, n_samples=512, n_speckle=100)
    a, b = w.rf_host(), w.rf_host()
    assert a.dtype == np.float32 and a.shape == (2, 128, 512) and np.array_equal(a, b)
    xs, zs, amp = w.scatterers()
    assert xs.size == w.n_points + w.n_speckle and np.all(amp[:w.n_points] == 10.0)
    assert np.all((xs >= -19e-3) & (xs <= 19e-3)) and np.all((zs >= 5e-3) & (zs <= 40e-3))


@pytest.mark.parametrize("n_el,kw,needle", [
    (300, {}, "at most 256 array elements"),
    (200, {"n_frames": 2}, "ensembles: at most 128 elements"),
    (160, {"line_scan": True}, "line scans: at most 128 elements"),
])
def test_wide_array_refusals_without_a_gpu(n_el, kw, needle):
    """Arrays beyond 128 elements take the one-launch forward kernel of single-frame
    plans (up to 256 elements); the other forms refuse them before any device work."""
    px = np.linspace(-1e-3, 1e-3, 4)
    extra = {}
    if kw.get("line_scan"):
        extra = dict(angles=np.zeros(px.size), virtual_sources=np.stack([px, np.full(px.size, 20e-3)], axis=1))
    P, keep, info = _params(n_el=n_el, px=px, **extra, **kw)
    L = _lib.load()
    h = C.c_void_p()
    st = L.fn_plan_pw_create(C.byref(P), C.byref(h))
    assert L.fn_status_id(st).decode() == "Unsupported"
    assert needle in L.fn_last_error().decode()


def test_noalloc_refuses_beyond_scratch_without_a_gpu():
    """fn_das_pw_noalloc: 75 angles over 1455 bins need 119 MB of spectra workspace,
    beyond its 32 MiB device array: refused before any device work."""
    from f_number_amd import synth
    from f_number_amd.das import pw_params

    w = synth.cpwc_full()
    P, keep, info = pw_params(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width,
                              w.pitch, w.c_0, w.f_bounds, 0.0, fa.windows.hann(), fa.f_numbers.grating.angle_lb(60, 1.5),
                              fa.normalizations.on())
    L = _lib.load()
    st = L.fn_das_pw_noalloc(C.byref(P), C.c_void_p(16), 0, C.c_void_p(16), None, None)
    assert L.fn_status_id(st).decode() == "Unsupported"
    assert "scratch" in L.fn_last_error().decode()
