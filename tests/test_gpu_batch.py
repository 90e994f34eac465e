"""GPU parity of the batching form: ensembles (frames x angles as transmit rows of
one pass, fnumber_hip.h fn_pw_params.n_frames).

Each frame of an ensemble must equal the single-frame plan of that frame's RF
(same kernels, other grouping: relative L2 <= 1e-5), and the FP64 oracle of
that frame (FP32 <= 1e-4; fp16 <= 1e-3, bf16x3 <= 1e-4)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    return torch


@pytest.fixture(scope="module")
def fa():
    import f_number_amd as fa

    fa.load()
    return fa


def _frames(n_frames, n_ang, n=20):
    from f_number_amd import synth

    ws, rfs = [], []
    for f in range(n_frames):
        w = synth.doppler_frame(f * 40)  # large steps so the frames differ visibly
        w.positions_x = np.linspace(w.positions_x[0], w.positions_x[-1], n)
        w.positions_z = np.linspace(w.positions_z[0], w.positions_z[-1], n + 4)
        w.n_samples = 1024
        w.n_speckle = 3000
        if n_ang != 5:
            w.angles_deg = np.linspace(-4.0, 4.0, n_ang) if n_ang > 1 else np.zeros(1)
        ws.append(w)
        rfs.append(w.rf_host())
    return ws, np.stack(rfs)


def _plan(fa, w, precision="fp32", n_frames=1, win="hann"):
    wo = {"hann": fa.windows.hann(), "tukey": fa.windows.tukey(0.5), "boxcar": fa.windows.boxcar()}[win]
    return fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=wo, F_number=fa.f_numbers.grating.angle_lb(60, 1.5),
                     precision=precision, output_scale="cpu", n_frames=n_frames)


@pytest.mark.parametrize("n_frames,n_ang,win", [(7, 5, "hann"), (9, 3, "tukey"), (20, 1, "boxcar"), (4, 11, "hann")])
def test_ensemble_frames_match_single_frame_plans(fa, torch, n_frames, n_ang, win):
    ws, rf = _frames(n_frames, n_ang)
    w = ws[0]
    print(f"case ensemble frames={n_frames} angles={n_ang} win={win}", flush=True)
    plan = _plan(fa, w, n_frames=n_frames, win=win)
    nx, nz = w.positions_x.size, w.positions_z.size
    imgs = torch.empty((n_frames, nx, nz), dtype=torch.complex64, device="cuda")
    plan.execute(torch.from_numpy(rf).cuda(), imgs)
    torch.cuda.synchronize()
    single = _plan(fa, w, win=win)
    one = torch.empty((nx, nz), dtype=torch.complex64, device="cuda")
    worst = 0.0
    for f in range(n_frames):
        single.execute(torch.from_numpy(rf[f]).cuda(), one)
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(imgs[f].cpu().numpy(), one.cpu().numpy()))
    print(f"  worst frame rel_l2 vs single-frame plans {worst:.2e}", flush=True)
    plan.close()
    single.close()
    assert worst <= 1e-5


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-4), ("fp16", 1e-3), ("bf16x3", 1e-4)])
def test_ensemble_matches_oracle(fa, oracle, torch, precision, tol):
    ws, rf = _frames(6, 5)
    w = ws[0]
    plan = _plan(fa, w, precision=precision, n_frames=6)
    imgs = torch.empty((6, w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    plan.execute(torch.from_numpy(rf).cuda(), imgs)
    torch.cuda.synchronize()
    plan.close()
    errs = []
    for f in (0, 3, 5):
        ref, _ = oracle.das_pw(w.positions_x, w.positions_z, rf[f], w.f_s, w.angles, w.width, w.pitch, w.c_0,
                               w.f_bounds, 0, ("hann", 0.0), oracle.fnumber("angle_lb", 60, 1.5, 10), "on")
        errs.append(rel_l2(imgs[f].cpu().numpy().T, ref))
    print(f"case ensemble {precision}: rel_l2 vs FP64 oracle {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= tol


def test_doppler_frames_differ_where_the_flow_is(fa, torch):
    """The flow phantom moves only the scatterers inside the tube: frame-to-frame
    differences concentrate around the tube (z = 20 mm +- 2 mm)."""
    from f_number_amd import synth

    w0, w1 = synth.doppler_frame(0), synth.doppler_frame(5)
    n = 64
    for w in (w0, w1):
        w.positions_x = np.linspace(-15e-3, 15e-3, n)
        w.positions_z = np.linspace(10e-3, 30e-3, n)
    rf = np.stack([w0.rf_host(), w1.rf_host()])
    plan = _plan(fa, w0, n_frames=2)
    imgs = torch.empty((2, n, n), dtype=torch.complex64, device="cuda")
    plan.execute(torch.from_numpy(rf).cuda(), imgs)
    torch.cuda.synchronize()
    plan.close()
    d = (imgs[1] - imgs[0]).abs().cpu().numpy()  # (x, z)
    z = w0.positions_z
    inside = np.abs(z - 20e-3) < 3e-3
    assert d[:, inside].mean() > 5 * d[:, ~inside].mean()
