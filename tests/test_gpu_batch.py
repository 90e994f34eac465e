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


@pytest.mark.parametrize("n_el,precision", [(64, "fp32"), (64, "fp16"), (40, "bf16x3"), (24, "fp32")])
def test_small_array_ensembles_match_single_frame_plans(fa, torch, n_el, precision):
    """Ensembles of arrays below 128 elements (16 and 8 element steps per lane, the 16-bit
    forms with one dynamic LDS tile): 5 frames x 3 angles, 8 frame slots per angle."""
    ws = []
    for f in range(5):
        w = synth_frame(f, n_el)
        ws.append(w)
    rf = np.stack([w.rf_host() for w in ws])
    w = ws[0]
    ens, single = _plan(fa, w, precision=precision, n_frames=5), _plan(fa, w, precision=precision)
    nx, nz = w.positions_x.size, w.positions_z.size
    imgs = torch.empty((5, nx, nz), dtype=torch.complex64, device="cuda")
    ens.execute(torch.from_numpy(rf).cuda(), imgs)
    one = torch.empty((nx, nz), dtype=torch.complex64, device="cuda")
    worst = 0.0
    for f in range(5):
        single.execute(torch.from_numpy(rf[f]).cuda(), one)
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(imgs[f].cpu().numpy(), one.cpu().numpy()))
    print(f"case {precision} ensemble at {n_el} elements: worst frame rel_l2 vs single-frame plans {worst:.2e}",
          flush=True)
    ens.close()
    single.close()
    assert worst <= 1e-5


def synth_frame(f, n_el):
    from f_number_amd import synth

    w = synth.scaled(synth.doppler_frame(f * 40), 20, 24, n_angles=3, n_samples=1024, n_speckle=3000,
                     n_elements=n_el)
    return w


@pytest.mark.parametrize("precision", ["fp32", "fp16x3"])
def test_ensemble_of_virtual_sources_matches_single_frame_plans(fa, torch, precision):
    """An ensemble of diverging-wave frames (virtual sources instead of steering angles):
    6 frames x 3 sources fill 8 frame slots per source of the angle-major rows, and
    every frame must equal its single-frame plan."""
    ws, rf = _frames(6, 3)
    w = ws[0]
    vs = [[-6e-3, -12e-3], [0.0, -10e-3], [5e-3, -14e-3]]

    def plan(n_frames):
        return fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width,
                         w.pitch, w.c_0, w.f_bounds, 0.0, window=fa.windows.hann(),
                         F_number=fa.f_numbers.grating.angle_lb(60, 1.5), precision=precision,
                         output_scale="cpu", n_frames=n_frames, virtual_sources=vs)

    ens, single = plan(6), plan(1)
    nx, nz = w.positions_x.size, w.positions_z.size
    imgs = torch.empty((6, nx, nz), dtype=torch.complex64, device="cuda")
    ens.execute(torch.from_numpy(rf).cuda(), imgs)
    one = torch.empty((nx, nz), dtype=torch.complex64, device="cuda")
    worst = 0.0
    for f in range(6):
        single.execute(torch.from_numpy(rf[f]).cuda(), one)
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(imgs[f].cpu().numpy(), one.cpu().numpy()))
    print(f"case {precision} virtual-source ensemble: worst frame rel_l2 vs single-frame plans {worst:.2e}",
          flush=True)
    ens.close()
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


@pytest.mark.parametrize("precision,max_mt", [("fp16", 8), ("fp32", 5), ("bf16x3", 4), ("fp16x3", 4)])
def test_ensemble_tile_groups_agree(fa, torch, precision, max_mt):
    """Ensembles take groups of up to 4 transmit tiles (FP32 at 128 elements up to 5, fp16 up to 8,
    one wave per SIMD from 6 on); every group size must give the image of the default
    grouping.  30 frames fill 32 frame slots of the angle-major rows."""
    ws, rf = _frames(30, 5)  # 5 angles x 32 slots = 160 operand rows = 10 tiles of 16
    w = ws[0]
    plan = _plan(fa, w, precision=precision, n_frames=30)
    shape = (30, w.positions_x.size, w.positions_z.size)
    ref = torch.empty(shape, dtype=torch.complex64, device="cuda")
    rf_d = torch.from_numpy(rf).cuda()
    plan.execute(rf_d, ref)
    split = plan.info()["bin_split"]
    for mt in range(1, max_mt + 1):
        plan.set_tiling(mt, split)
        img = torch.empty_like(ref)
        plan.execute(rf_d, img)
        torch.cuda.synchronize()
        err = (torch.linalg.vector_norm(img - ref) / torch.linalg.vector_norm(ref)).item()
        print(f"case {precision} ensemble m_tiles {mt}: rel_l2 vs default grouping {err:.2e}", flush=True)
        assert err < 1e-6
    with pytest.raises(fa.FNumberError):
        plan.set_tiling(max_mt + 1, split)
    plan.close()


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


def _cpwc(n_x, n_z, n_ang, seed, n_el=128, n_samples=1024):
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), n_x, n_z, n_samples=n_samples, n_speckle=1500, n_elements=n_el)
    w.angles_deg = np.linspace(-10.0, 10.0, n_ang) if n_ang > 1 else np.zeros(1)
    w.seed = seed
    return w


@pytest.mark.parametrize("n_ang", [11, 40])
def test_autotune_keeps_the_image(fa, torch, n_ang):
    w = _cpwc(48, 40, n_ang, 31 + n_ang)
    rf = w.rf_device("cuda:0")
    plan = _plan(fa, w)
    a = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    plan.execute(rf, a)
    res = plan.autotune(rf, b, repetitions=2)
    plan.execute(rf, b)
    torch.cuda.synchronize()
    info = plan.info()
    print(f"case autotune n_ang={n_ang}: {res} -> m_tiles {info['m_tiles']} bin_split {info['bin_split']}", flush=True)
    assert res["n_candidates"] >= 2 and res["best_ms"] <= res["default_ms"] * (1 + 1e-9)
    assert info["m_tiles"] == res["m_tiles"] and info["bin_split"] == res["bin_split"]
    assert (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a)).item() < 1e-6
    with pytest.raises(fa.FNumberError):
        plan.set_tiling(9, 1)
    plan.close()


def test_batch_graph_replays_match_eager(fa, torch):
    """One graph over heterogeneous plans (grids, angle counts, precisions, an ensemble
    and a synthetic-aperture plan); every replay equals the eager images, with the
    program's own eager work in between; new buffers trigger a new capture."""
    from f_number_amd import synth

    jobs = []
    for (nx, nz, na, prec) in [(48, 40, 11, "fp32"), (32, 64, 1, "fp16"), (64, 32, 31, "bf16x3"), (40, 40, 3, "fp32")]:
        w = _cpwc(nx, nz, na, 100 + nx + na)
        jobs.append((_plan(fa, w, precision=prec), w.rf_device("cuda:0"), (nx, nz), 1))
    ws, rfe = _frames(5, 5)
    jobs.append((_plan(fa, ws[0], precision="fp16", n_frames=5), torch.from_numpy(rfe).cuda(),
                 (ws[0].positions_x.size, ws[0].positions_z.size), 5))
    ws = synth.scaled(synth.saft_full(), 20, 16, n_samples=768, n_speckle=300, n_elements=32)
    ws.pitch = 4 * 0.3048e-3
    ws.width = 4 * 0.27e-3
    sa = fa.SaPlan(ws.positions_x, ws.positions_z, ws.n_samples, ws.n_elements, ws.f_s, ws.width, ws.pitch, ws.c_0,
                   ws.f_bounds, 0.0, output_scale="cpu")
    jobs.append((sa, ws.rf_device("cuda:0"), (20, 16), 1))
    eager, outs = [], []
    for plan, rf, (nx, nz), nfr in jobs:
        shape = (nfr, nx, nz) if nfr > 1 else (nx, nz)
        e = torch.empty(shape, dtype=torch.complex64, device="cuda")
        plan.execute(rf, e)
        eager.append(e)
        outs.append(torch.empty_like(e))
    torch.cuda.synchronize()
    batch = fa.Batch([j[0] for j in jobs], n_streams=3)
    rfs = [j[1] for j in jobs]
    for replay in range(4):
        for o in outs:
            o.fill_(float("nan"))  # eager work between replays
        batch.execute(rfs, outs)
        torch.cuda.synchronize()
        for i, (o, e) in enumerate(zip(outs, eager)):
            err = (torch.linalg.vector_norm(o - e) / torch.linalg.vector_norm(e)).item()
            assert err < 1e-6, f"replay {replay} job {i}: {err}"
    info = batch.info()
    print(f"case batch: {info}", flush=True)
    assert info["captures"] == 1 and info["graph_nodes"] >= len(jobs) * 2
    outs[0] = torch.empty_like(outs[0])
    batch.execute(rfs, outs)
    torch.cuda.synchronize()
    assert batch.info()["captures"] == 2
    assert (torch.linalg.vector_norm(outs[0] - eager[0]) / torch.linalg.vector_norm(eager[0])).item() < 1e-6
    batch.close()
    for j in jobs:
        j[0].close()


def test_persistent_batch_matches_eager(fa, torch):
    """Plans sharing a kernel class (uneven grids, angle counts and band splits) run as
    one persistent launch over their work items: every replay equals the eager images,
    with persistent launches on and off, and the persistent graph has fewer nodes."""
    from f_number_amd import synth

    specs = [  # (n_x, n_z, angles, precision, (m_tiles, bin_split) or None = default tiling)
        (48, 40, 11, "fp32", (1, 1)), (32, 64, 1, "fp32", (1, 4)), (64, 32, 3, "fp32", (1, 2)),
        (40, 40, 11, "fp32", None), (130, 24, 5, "fp32", None),
        (28, 36, 24, "fp32", (1, 3)),  # two transmit groups of one tile each
        (36, 28, 31, "fp32", None), (44, 20, 17, "fp32", None),  # two tiles
        (48, 40, 11, "fp16", None), (56, 24, 3, "fp16", (1, 2)), (24, 56, 1, "fp16", None)]
    jobs = []
    for k, (nx, nz, na, prec, tiling) in enumerate(specs):
        w = _cpwc(nx, nz, na, 500 + k)
        plan = _plan(fa, w, precision=prec)
        if tiling is not None:
            plan.set_tiling(*tiling)
        jobs.append((plan, w.rf_device("cuda:0"), (nx, nz)))
    for nx, nz in [(20, 16), (28, 12)]:
        ws = synth.scaled(synth.saft_full(), nx, nz, n_samples=768, n_speckle=300, n_elements=32)
        ws.pitch = 4 * 0.3048e-3
        ws.width = 4 * 0.27e-3
        sa = fa.SaPlan(ws.positions_x, ws.positions_z, ws.n_samples, ws.n_elements, ws.f_s, ws.width, ws.pitch,
                       ws.c_0, ws.f_bounds, 0.0, output_scale="cpu")
        jobs.append((sa, ws.rf_device("cuda:0"), (nx, nz)))
    eager, outs = [], []
    for plan, rf, shape in jobs:
        e = torch.empty(shape, dtype=torch.complex64, device="cuda")
        plan.execute(rf, e)
        eager.append(e)
        outs.append(torch.empty_like(e))
    torch.cuda.synchronize()
    rfs = [j[1] for j in jobs]
    batch = fa.Batch([j[0] for j in jobs], n_streams=3, persistent=True)
    nodes = {}
    for persistent in (True, False):
        batch.set_persistent(persistent)
        print(f"case persistent={persistent}", flush=True)
        for replay in range(3):
            for o in outs:
                o.fill_(float("nan"))  # eager work between replays
            batch.execute(rfs, outs)
            torch.cuda.synchronize()
            for i, (o, e) in enumerate(zip(outs, eager)):
                err = (torch.linalg.vector_norm(o - e) / torch.linalg.vector_norm(e)).item()
                assert err < 1e-6, f"persistent={persistent} replay {replay} job {i}: {err}"
        nodes[persistent] = batch.info()["graph_nodes"]
    print(f"case graph nodes: {nodes}", flush=True)
    assert nodes[True] < nodes[False]
    batch.close()
    for j in jobs:
        j[0].close()


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("fp16", 1e-4), ("bf16x3", 1e-5), ("fp16x3", 1e-5)])
def test_row_stacks_match_eager(fa, torch, oracle, precision, tol):
    """Plans that differ only in steering angles and RF (one grid, 1 to 31 angles, angle
    counts that do not fill a tile) beamformed as one row stack: every plan's image
    equals its own eager execute on every replay, with eager work in between, and the
    first plan's the FP64 oracle.  Plans of another grid, another window or more than
    48 rows stay on their own; with stacking off the batch gives the same images."""
    specs = [(40, 36, 1, "hann"), (40, 36, 3, "hann"), (40, 36, 11, "hann"), (40, 36, 31, "hann"),
             (40, 36, 5, "hann"), (40, 36, 7, "hann"), (40, 36, 11, "hann"),
             (40, 36, 3, "tukey"), (40, 36, 5, "tukey"),  # a second stack (other window)
             (32, 36, 3, "hann"),                         # another grid: on its own
             (40, 36, 60, "hann")]                        # four tiles of rows: on its own
    jobs = []
    for k, (nx, nz, na, win) in enumerate(specs):
        w = _cpwc(nx, nz, na, 900 + k)
        jobs.append((_plan(fa, w, precision=precision, win=win), w, w.rf_device("cuda:0"), (nx, nz)))
    eager, outs = [], []
    for plan, w, rf, shape in jobs:
        e = torch.empty(shape, dtype=torch.complex64, device="cuda")
        plan.execute(rf, e)
        eager.append(e)
        outs.append(torch.empty_like(e))
    torch.cuda.synchronize()
    rfs = [j[2] for j in jobs]
    batch = fa.Batch([j[0] for j in jobs], n_streams=3)
    for stacking in (True, False):
        batch.set_stacking(stacking)
        for replay in range(3):
            for o in outs:
                o.fill_(float("nan"))  # eager work between replays
            batch.execute(rfs, outs)
            torch.cuda.synchronize()
            for i, (o, e) in enumerate(zip(outs, eager)):
                err = (torch.linalg.vector_norm(o - e) / torch.linalg.vector_norm(e)).item()
                assert err < tol, f"stacking={stacking} replay {replay} job {i}: {err}"
        info = batch.info()
        print(f"case {precision} stacking={stacking}: {info}", flush=True)
        if stacking:
            assert info["row_stacks"] == 2 and info["stacked_plans"] == 9
        else:
            assert info["row_stacks"] == 0
    w = jobs[2][1]
    ref, _ = oracle.das_pw(w.positions_x, w.positions_z, w.rf_host(), w.f_s, w.angles, w.width, w.pitch, w.c_0,
                           w.f_bounds, 0, ("hann", 0.0), oracle.fnumber("angle_lb", 60, 1.5, 10), "on")
    batch.set_stacking(True)
    batch.execute(rfs, outs)
    torch.cuda.synchronize()
    e = rel_l2(outs[2].cpu().numpy().T, ref)
    print(f"  stacked plan 2 vs oracle: {e:.3e}", flush=True)
    assert e <= {"fp32": 1e-4, "fp16": 1e-3, "bf16x3": 1e-4, "fp16x3": 1e-4}[precision]
    batch.close()
    for j in jobs:
        j[0].close()
