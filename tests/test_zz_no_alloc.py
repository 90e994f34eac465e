"""fn_das_pw_noalloc (the entry point that allocates nothing) against the plan
entry point fn_plan_execute on the same device buffers: two calls in a row, the
second with a smaller input (fewer voxels, angles and samples, another window and
semantics), and the refusal of a call that needs more scratch than the library holds.

Bar: the image within rel-L2 1e-5 of the plan's (the same kernels; the bin-chunk
partial images are summed with float atomics, in either order) and within the
1e-4 / -80 dB parity bar of the FP64 oracle; the weights within 1e-6.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import peak_db, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import f_number_amd as fa

    fa.load()
    return fa


def _call_noalloc(fa, P, rf, img, w):
    import torch

    from f_number_amd import _lib

    s = torch.cuda.current_stream(rf.device).cuda_stream
    return _lib.load().fn_das_pw_noalloc(C.byref(P), C.c_void_p(rf.data_ptr()), _lib.DTYPE["float32"],
                                         C.c_void_p(img.data_ptr()), C.c_void_p(w.data_ptr()), C.c_void_p(s))


def _case(fa, oracle, w, win, win_o, semantics, precision):
    import torch

    from f_number_amd import _lib
    from f_number_amd.das import pw_params

    rf_h = w.rf_host()
    rf = torch.from_numpy(rf_h).cuda()
    F = fa.f_numbers.grating.angle_lb(60, 1.5)
    args = (w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch, w.c_0,
            w.f_bounds, 0.0, win, F, fa.normalizations.on())
    P, keep, info = pw_params(*args[:11], win, F, fa.normalizations.on(), 0, semantics, precision, "cpu")
    nx, nz = w.positions_x.size, w.positions_z.size
    print(f"case noalloc {nz}x{nx} angles={w.angles.size} n_t={w.n_samples} {semantics} {precision}", flush=True)
    img = torch.full((nx, nz), complex(7.0, -7.0), dtype=torch.complex64, device="cuda")  # stale contents
    wt = torch.full((nx, nz), 3.0, dtype=torch.float32, device="cuda")
    _lib.check(_call_noalloc(fa, P, rf, img, wt), "gpu_bf_das_pw_rf")
    plan = fa.PwPlan(*args, semantics=semantics, precision=precision, output_scale="cpu")
    img2 = torch.empty((nx, nz), dtype=torch.complex64, device="cuda")
    wt2 = torch.empty((nx, nz), dtype=torch.float32, device="cuda")
    plan.execute(rf, img2, wt2)
    torch.cuda.synchronize()
    a, b = img.cpu().numpy().T, img2.cpu().numpy().T
    e_plan = rel_l2(a, b)
    e_w = float(np.max(np.abs(wt.cpu().numpy() - wt2.cpu().numpy())) / np.max(np.abs(wt2.cpu().numpy())))
    ref, _ = oracle.das_pw(w.positions_x, w.positions_z, rf_h, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds,
                           0, win_o, oracle.fnumber("angle_lb", 60, 1.5, 10), "on", semantics=semantics)
    e_o, d_o = rel_l2(a, ref), peak_db(a, ref)
    print(f"  vs plan rel_l2={e_plan:.2e} weights={e_w:.2e}; vs oracle rel_l2={e_o:.2e} peak={d_o:.1f} dB",
          flush=True)
    plan.close()
    assert e_plan <= 1e-5 and e_w <= 1e-6
    assert e_o <= 1e-4 and d_o <= -80.0


def test_noalloc_matches_plan_two_calls(fa, oracle):
    from f_number_amd import synth

    w1 = synth.scaled(synth.cpwc_small(), 96, 80, n_angles=11, n_samples=1024, n_speckle=1500)
    _case(fa, oracle, w1, fa.windows.tukey(0.5), ("tukey", 0.5), "cpu", "fp32")
    # second call, smaller: the scratch keeps the first call's contents
    w2 = synth.scaled(synth.cpwc_small(), 40, 36, n_angles=3, n_samples=768, n_speckle=600)
    _case(fa, oracle, w2, fa.windows.hann(), ("hann", 0.0), "mex", "fp32")


def test_noalloc_refuses_beyond_scratch(fa):
    """75 angles over 1455 bins need 119 MB of spectra: refused before any device work."""
    import torch

    from f_number_amd import _lib, synth
    from f_number_amd.das import pw_params

    w = synth.cpwc_full()
    P, keep, info = pw_params(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width,
                              w.pitch, w.c_0, w.f_bounds, 0.0, fa.windows.hann(), fa.f_numbers.grating.angle_lb(60, 1.5),
                              fa.normalizations.on())
    dummy = torch.empty(4, dtype=torch.float32, device="cuda")
    st = _call_noalloc(fa, P, dummy, dummy, dummy)
    L = _lib.load()
    assert L.fn_status_id(st).decode() == "Unsupported"
    assert "scratch" in L.fn_last_error().decode()
