"""Throughput of the Fourier-domain DAS beamformer on one MI355X.

Workload: plain CPWC at full size -- 128 elements, 75 steered plane waves
(-18.5..18.5 deg, 0.5 deg steps), 3328 samples, 1-9 MHz band, 1024 x 512
voxels -- with the frequency-dependent F-number grating.angle_lb(60, 1.5) and
a Hann receive window (headline), plus the fixed F = 1.5 boxcar arm.  One step
is one compounded frame: band-limited spectra of all 9600 channels plus the
fused beamforming kernel, from device-resident float32 RF to a device-resident
complex64 image.  Synthetic seeded data (the reference ships none).

Prints one JSON line: Mvoxel*frames/s, the roofline of the dominant kernel
(beamforming, MFMA-bound, FP32 peak) from HIP events recorded inside the
library on the launch stream, and the reference CPU path (FP64 C restatement
of das_pw.m, oracle/) timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 matrix
PEAK_16BIT_TFLOPS = 2500.0  # MI355X dense BF16 / FP16 matrix
PEAK_HBM_GBS = 8000.0
CPU_LABEL = "reference CPU path (FP64 C restatement of das_pw.m / das_sa.m; MATLAB unavailable)"


def host_cpus():
    """(threads for the CPU baseline, CPUs this process may run on, CPU model): every CPU
    of the process's affinity set, or OMP_NUM_THREADS when that is smaller (the share
    of a shared host that the process is given)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    n = avail
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, avail, model


def precision_roofline(prec, flops, beamform_ms):
    """Algorithmic and executed MFMA fractions of a precision arm's beamforming kernel:
    algorithmic = 8 N_f N_vox N_el N_tx / time / peak of the contraction precision;
    executed counts the MFMA work the form performs (FP32: Gauss's 3 real products per
    complex MAC, 3/4 of the dense 4; split 16-bit forms: 4 real products x 3 passes; fp16: 4)."""
    peak = PEAK_FP32_TFLOPS if prec == "fp32" else PEAK_16BIT_TFLOPS
    executed_per_alg = {"fp32": 0.75, "bf16x3": 3.0, "fp16x3": 3.0, "fp16": 1.0}[prec]
    achieved = flops / (beamform_ms / 1e3) / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "executed_frac": achieved * executed_per_alg / peak}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--arm", choices=["both", "hann", "boxcar"], default="both")
    ap.add_argument("--workload", choices=["cpwc_full", "cpwc_small"], default="cpwc_full")
    ap.add_argument("--precisions", default="bf16x3,fp16x3,fp16",
                    help="mixed-precision arms of the headline configuration (comma list, '' for none)")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation-form measurements")
    ap.add_argument("--no-batching", action="store_true", help="skip the batching-form measurements")
    ap.add_argument("--no-saft", action="store_true", help="skip the SAFT precision-form measurements")
    ap.add_argument("--no-line-scan", action="store_true", help="skip the line-by-line scanning measurement")
    ap.add_argument("--no-wide", action="store_true", help="skip the 192 / 256-element arrays")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the 11-angle 256 x 256 CPWC arm (the CPU-runnable workload at GPU scale)")
    ap.add_argument("--bin-split", type=int, default=0,
                    help="bin split of the headline arm's plan (0: the plan's default tiling)")
    ap.add_argument("--ensemble-frames", type=int, default=400)
    ap.add_argument("--queue-jobs", type=int, default=64)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    return ap.parse_args()


def measured_traffic(workload):
    """HBM bytes per launch of the beamforming kernel from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (each scaled by its reading for a device copy of known size),
    kept in profiles/beamform_traffic.json; None when that workload was not profiled."""
    path = os.path.join(ROOT, "profiles", "beamform_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)[workload]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def spectra_roofline(w, n_f, spectra_ms):
    """The band-limited DFT stage (spectra_kernel): its algorithmic HBM bytes -- the RF
    read once (4 N_t bytes per channel) and the band spectra written once (8 N_f bytes per
    channel) -- over its event-timed duration against 8 TB/s, and its DFT arithmetic
    (4 flop per sample and bin of a real channel) against the FP32 matrix peak, on which
    it runs."""
    n_ch = w.n_tx * w.n_elements
    bytes_alg = n_ch * (4.0 * w.n_samples + 8.0 * n_f)
    flops = 4.0 * n_ch * w.n_samples * n_f
    gbs = bytes_alg / (spectra_ms / 1e3) / 1e9
    tfs = flops / (spectra_ms / 1e3) / 1e12
    return {"kernel": "spectra_kernel", "bytes_per_launch": bytes_alg, "achieved_gbs": gbs,
            "hbm_frac": gbs / PEAK_HBM_GBS, "flops_per_launch": flops, "achieved_tflops": tfs,
            "mfma_frac": tfs / PEAK_FP32_TFLOPS, "avg_launch_ms": spectra_ms,
            "bound": "mfma" if tfs / PEAK_FP32_TFLOPS > gbs / PEAK_HBM_GBS else "hbm"}


def sample_column(w):
    """lateral column of the bounded CPU sample (the centre one)"""
    return w.positions_x.size // 2


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cpu_baseline(w, window_name, F_kind, F_params, n_threads):
    """Reference CPU path (FP64 C restatement of das_pw.m:302-485) on a bounded sample:
    all spectra + one lateral column of voxels, extrapolated linearly in voxels.  Also
    returns the FP64 image of that column (CPU output scale), against which the GPU
    arms' errors are stated."""
    from oracle import oracle as O

    O.build()
    rf = w.rf_host()
    F = O.fnumber(F_kind, *F_params)
    ix_col = sample_column(w)
    sel = (np.full(w.positions_z.size, ix_col), np.arange(w.positions_z.size))
    # spectra alone (the FFT of das_pw.m:383)
    N, lb, ub = O.grid(w.n_samples, w.n_elements, w.pitch, w.positions_z[0], w.f_s, w.c_0, *w.f_bounds)
    t0 = time.perf_counter()
    O.spectra(rf.reshape(-1, w.n_samples), N, lb, ub, mode=1, n_threads=n_threads)
    t_spec = time.perf_counter() - t0
    t0 = time.perf_counter()
    col, _ = O.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0,
                      (window_name, 0.0), F, "on", n_threads=n_threads, sel=sel)
    t_col = time.perf_counter() - t0 - t_spec  # das_pw recomputes the spectra
    t_frame = t_spec + max(t_col, 0.0) * w.positions_x.size
    # one thread: the voxel loop on 32 voxels spread down the column (its spectra stage
    # on n_threads, timed above and subtracted) and the spectra of 1/16 of the channels
    n1 = 32
    iz1 = np.linspace(0, w.positions_z.size - 1, n1).round().astype(int)
    t0 = time.perf_counter()
    O.das_pw(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0,
             (window_name, 0.0), F, "on", n_threads=1, n_threads_spectra=n_threads,
             sel=(np.full(n1, ix_col), iz1))
    t_vox1 = max(time.perf_counter() - t0 - t_spec, 0.0) / n1
    ch = rf.reshape(-1, w.n_samples)
    sub = ch[: max(1, ch.shape[0] // 16)]
    t0 = time.perf_counter()
    O.spectra(sub, N, lb, ub, mode=1, n_threads=1)
    t_spec1 = (time.perf_counter() - t0) * ch.shape[0] / sub.shape[0]
    t_frame1 = t_spec1 + t_vox1 * w.n_voxels
    _, avail, model = host_cpus()
    return {
        "value": w.n_voxels / t_frame / 1e6,
        "unit": "Mvoxel*frames/s",
        "cores": n_threads,
        "host_cpus": avail,
        "cpu_model": model,
        "label": CPU_LABEL,
        "kind": "port",
        "sample": f"spectra of all {w.n_tx * w.n_elements} channels + 1 of {w.positions_x.size} lateral "
                  f"columns ({w.positions_z.size} voxels, all {w.n_tx} angles), extrapolated x{w.positions_x.size}; "
                  f"sample wall {t_spec + t_col:.1f} s",
        "single_thread": {"value": w.n_voxels / t_frame1 / 1e6, "cores": 1,
                          "sample": f"voxel loop on {n1} voxels of that column and spectra of {sub.shape[0]} of "
                                    f"{ch.shape[0]} channels, one thread each, extrapolated to the frame"},
    }, np.asarray(col).reshape(-1)


def saft_fp64_check(w, rf, kw_oracle, n_threads, gpu_cols, n_dft):
    """FP64 oracle (das_sa.m:289-509 restated) on the centre lateral column of the
    SAFT workload; rel-L2 of each GPU arm's column against it, and the oracle's own
    time as the SAFT CPU baseline (spectra + one column, extrapolated)."""
    from oracle import oracle as O

    O.build()
    ix_col = sample_column(w)
    sel = (np.full(w.positions_z.size, ix_col), np.arange(w.positions_z.size))
    N, lb, ub = O.grid(w.n_samples, w.n_elements, w.pitch, w.positions_z[0], w.f_s, w.c_0, *w.f_bounds)
    t0 = time.perf_counter()
    O.spectra(rf.reshape(-1, w.n_samples), N, lb, ub, mode=1, n_threads=n_threads)
    t_spec = time.perf_counter() - t0
    t0 = time.perf_counter()
    col, _ = O.das_sa(w.positions_x, w.positions_z, rf, w.f_s, w.width, w.pitch, w.c_0, w.f_bounds, 0.0,
                         n_threads=n_threads, sel=sel, **kw_oracle)
    t_col = max(time.perf_counter() - t0 - t_spec, 0.0)
    col = np.asarray(col).reshape(-1)
    errs = {k: rel_l2(v * n_dft, col) for k, v in gpu_cols.items()}  # GPU arms: MEX scale 2/N, oracle: 2
    t_frame = t_spec + t_col * w.positions_x.size
    _, avail, model = host_cpus()
    return errs, {"value": w.n_voxels / t_frame / 1e6, "unit": "Mvoxel*frames/s", "cores": n_threads,
                  "host_cpus": avail, "cpu_model": model, "label": CPU_LABEL, "kind": "port",
                  "sample": f"spectra of all {w.n_elements ** 2} channels + 1 of {w.positions_x.size} lateral columns "
                            f"({w.positions_z.size} voxels, 128 transmits), extrapolated x{w.positions_x.size}; "
                            f"sample wall {t_spec + t_col:.1f} s"}


def run_saft(steps, warmup, dev, precs):
    """SAFT precision form: full synthetic aperture 128 x 128 x 3000 samples, 512 x 512
    voxels, Tukey(0.5) transmit and receive windows, angle_lb(60, 1.5) F-numbers on both
    apertures, normalization on; FP32 and the mixed-precision contractions (TOF and the
    phase reduction in FP64 in all of them).  Returns timings and the centre column of
    every arm's image for the FP64 check."""
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    w = synth.saft_full()
    rf = w.rf_device(dev)
    kw = dict(window_tx=fa.windows.tukey(0.5), window_rx=fa.windows.tukey(0.5),
              F_number_tx=fa.f_numbers.grating.angle_lb(60, 1.5), F_number_rx=fa.f_numbers.grating.angle_lb(60, 1.5),
              normalization_tx=fa.normalizations.on(), normalization_rx=fa.normalizations.on())
    res, cols = {}, {}
    ix = sample_column(w)
    ref = None
    for p in ["fp32"] + [q for q in precs if q != "fp32"]:
        plan = fa.SaPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.f_s, w.width, w.pitch, w.c_0,
                         w.f_bounds, 0.0, precision=p, index_gpu=dev.index or 0, **kw)
        info = plan.info()
        img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device=dev)
        stream = torch.cuda.current_stream(dev)
        for _ in range(warmup):
            plan.execute(rf, img, stream=stream)
        torch.cuda.synchronize(dev)
        plan.set_profiling(True)
        plan.get_profile()
        t0 = time.perf_counter()
        for _ in range(steps):
            plan.execute(rf, img, stream=stream)
        torch.cuda.synchronize(dev)
        wall = (time.perf_counter() - t0) / steps
        spec_ms, beam_ms, n = plan.get_profile()
        plan.close()
        beam_ms /= max(n, 1)
        r = {"ms_per_step": wall * 1e3, "beamform_ms": beam_ms, "prep_ms": spec_ms / max(n, 1),
             "value": w.n_voxels / wall / 1e6,
             "achieved_tflops": info["useful_flops_per_frame"] / (beam_ms / 1e3) / 1e12,
             "roofline": precision_roofline(p, info["useful_flops_per_frame"], beam_ms),
             "finite": bool(torch.isfinite(torch.view_as_real(img)).all().item())}
        if p == "fp32":
            ref = img.clone()
        else:
            r["rel_l2_vs_fp32"] = (torch.linalg.vector_norm(img - ref) / torch.linalg.vector_norm(ref)).item()
            r["beamform_speedup_vs_fp32"] = res["fp32"]["beamform_ms"] / beam_ms
        res[p] = r
        cols[p] = img[ix].cpu().numpy().astype(np.complex128)
        del img
    rf_host = rf.cpu().numpy()  # the very input of the GPU arms, for the FP64 check
    return {"workload": w.name, "n_voxels": int(w.n_voxels), "N_dft": info["N_dft"], "n_f": info["n_f"],
            "flops_per_frame": info["useful_flops_per_frame"], "flops_dense_per_frame": info["flops_per_frame"],
            "arms": res}, w, cols, rf_host


def run_arm(w, plan_kw, steps, warmup, dev, keep_image=False, bin_split=0):
    import torch

    import f_number_amd as fa

    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, index_gpu=dev.index or 0, **plan_kw)
    if bin_split > 0:
        plan.set_tiling(plan.info()["m_tiles"], bin_split)
    info = plan.info()
    rf = w.rf_device(dev)
    img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)
    for _ in range(warmup):
        plan.execute(rf, img, stream=stream)
    torch.cuda.synchronize(dev)
    plan.set_profiling(True)
    plan.get_profile()  # reset
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute(rf, img, stream=stream)
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t0
    spec_ms, beam_ms, n = plan.get_profile()
    plan.set_profiling(False)
    ok = bool(torch.isfinite(torch.view_as_real(img)).all().item())
    plan.close()
    return {
        "image": img if keep_image else None,
        "ms_per_step": wall / steps * 1e3,
        "spectra_ms": spec_ms / max(n, 1),
        "beamform_ms": beam_ms / max(n, 1),
        "info": info,
        "finite": ok,
    }


def _time_ms(fn, steps, warmup, dev):
    """mean ms per call of fn() from events recorded on the current stream"""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize(dev)
    return e0.elapsed_time(e1) / steps


def run_small(steps, warmup, dev):
    """The 11-angle CPWC workload of the CPU-runnable case at GPU scale: 128 elements,
    -10..10 deg, 2048 samples, 2.25-6.75 MHz, 256 x 256 voxels, grating.angle_lb(60, 1.5)
    with Tukey(0.5), normalization on; the default tiling and the autotuned one."""
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    w = synth.cpwc_small()
    kw = dict(window=fa.windows.tukey(0.5), F_number=fa.f_numbers.grating.angle_lb(60, 1.5))
    r = run_arm(w, kw, max(steps, 10), warmup, dev)
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, index_gpu=dev.index or 0, **kw)
    rf = w.rf_device(dev)
    img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device=dev)
    tune = plan.autotune(rf, img, repetitions=3)
    plan.set_profiling(True)
    plan.get_profile()
    for _ in range(max(steps, 10)):
        plan.execute(rf, img)
    torch.cuda.synchronize(dev)
    _, beam_tuned, n = plan.get_profile()
    plan.close()
    beam_tuned /= max(n, 1)
    flops = r["info"]["useful_flops_per_frame"]
    return {"workload": w.name, "n_voxels": int(w.n_voxels), "value": w.n_voxels / (r["ms_per_step"] / 1e3) / 1e6,
            "unit": "Mvoxel*frames/s", "ms_per_step": r["ms_per_step"], "spectra_ms": r["spectra_ms"],
            "beamform_ms": r["beamform_ms"], "roofline": precision_roofline("fp32", flops, r["beamform_ms"]),
            "tiling": {"m_tiles": r["info"]["m_tiles"], "bin_split": r["info"]["bin_split"]},
            "autotuned": {"m_tiles": tune["m_tiles"], "bin_split": tune["bin_split"], "beamform_ms": beam_tuned,
                          "roofline": precision_roofline("fp32", flops, beam_tuned)}}


def run_wide(steps, warmup, dev):
    """Arrays beyond 128 elements: the 11-angle CPWC workload (2048 samples, 256 x 256,
    Tukey(0.5), angle_lb(60, 1.5)) at 192 and 256 elements, and the full synthetic aperture
    at 192 and 256 elements (3000 samples, 256 x 256); FP32 and bf16x3, default tilings."""
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    out = {}
    for n_el in (192, 256):
        w = synth.scaled(synth.cpwc_small(), 256, 256, n_elements=n_el)
        w.seed = 20260100 + n_el
        for prec in ("fp32", "bf16x3"):
            kw = dict(window=fa.windows.tukey(0.5), F_number=fa.f_numbers.grating.angle_lb(60, 1.5), precision=prec)
            r = run_arm(w, kw, steps, warmup, dev)
            flops = r["info"]["useful_flops_per_frame"]
            out[f"pw_{n_el}el_{prec}"] = {"beamform_ms": r["beamform_ms"], "spectra_ms": r["spectra_ms"],
                                          "value": w.n_voxels / (r["ms_per_step"] / 1e3) / 1e6,
                                          "m_tiles": r["info"]["m_tiles"], "bin_split": r["info"]["bin_split"],
                                          "roofline": precision_roofline(prec, flops, r["beamform_ms"]),
                                          "finite": r["finite"]}
        ws = synth.scaled(synth.saft_full(), 256, 256, n_elements=n_el)
        ws.seed = 20260200 + n_el
        rf = ws.rf_device(dev)
        for prec in ("fp32", "bf16x3"):
            plan = fa.SaPlan(ws.positions_x, ws.positions_z, ws.n_samples, ws.n_elements, ws.f_s, ws.width, ws.pitch,
                             ws.c_0, ws.f_bounds, 0.0, window_tx=fa.windows.tukey(0.5),
                             window_rx=fa.windows.tukey(0.5), F_number_tx=fa.f_numbers.grating.angle_lb(60, 1.5),
                             F_number_rx=fa.f_numbers.grating.angle_lb(60, 1.5), precision=prec,
                             index_gpu=dev.index or 0)
            info = plan.info()
            img = torch.empty((ws.positions_x.size, ws.positions_z.size), dtype=torch.complex64, device=dev)
            for _ in range(warmup):
                plan.execute(rf, img)
            torch.cuda.synchronize(dev)
            plan.set_profiling(True)
            plan.get_profile()
            for _ in range(steps):
                plan.execute(rf, img)
            torch.cuda.synchronize(dev)
            spec_ms, beam_ms, n = plan.get_profile()
            plan.close()
            beam_ms /= max(n, 1)
            out[f"sa_{n_el}el_{prec}"] = {"beamform_ms": beam_ms, "prep_ms": spec_ms / max(n, 1),
                                          "roofline": precision_roofline(prec, info["useful_flops_per_frame"], beam_ms),
                                          "finite": bool(torch.isfinite(torch.view_as_real(img)).all().item())}
        del rf
    return out


def run_diff(steps, warmup, dev):
    """Differentiation form on the layered-medium acquisition (11 angles, 2048 samples,
    384 x 384): forward, adjoint, tangent and vector-Jacobian timings (all enqueued on
    torch's current stream; NULL stream handle = the default stream) and a 50-iteration
    autofocus from c = 1480 m/s."""
    import numpy as np
    import torch

    import f_number_amd as fa
    from f_number_amd import autofocus as AF
    from f_number_amd import synth

    w = synth.autofocus_layer()
    rf = w.rf_device(dev)
    kw = dict(window=fa.windows.tukey(0.5), F_number=fa.f_numbers.grating.angle_lb(60, 1.5),
              normalization=fa.normalizations.on())
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     1480.0, w.f_bounds, 0.0, **kw)
    n = (w.positions_x.size, w.positions_z.size)
    img = torch.empty(n, dtype=torch.complex64, device=dev)
    dimg = torch.empty_like(img)
    adj = torch.empty((w.n_tx, w.n_elements, w.n_samples), dtype=torch.float32, device=dev)
    plan.set_tof(1500.0, w.delay_offsets * 0.5)
    t_fwd = _time_ms(lambda: plan.execute(rf, img), steps, warmup, dev)
    y = img.clone()
    # first adjoint of the plan: also builds its per-plan tables (phase words for this
    # time of flight, the normalization s / w_n(v)); later calls reuse them
    t_first = _time_ms(lambda: plan.adjoint(y, adj), 1, 0, dev)
    t_adj = _time_ms(lambda: plan.adjoint(y, adj), steps, warmup, dev)
    t_tan = _time_ms(lambda: plan.execute_tangent(rf, img, dimg), steps, warmup, dev)
    g = (2 * img).contiguous()
    t_vjp = _time_ms(lambda: plan.vjp(rf, g, dimg), steps, warmup, dev)
    # the forward at its autotuned tiling, against which the adjoint is also compared
    tune = plan.autotune(rf, img, repetitions=3)
    t_fwd_tuned = _time_ms(lambda: plan.execute(rf, img), steps, warmup, dev)
    res = AF.autofocus(plan, rf, n_iter=50, c0=1480.0)
    plan.close()
    truth = w.delay_offsets - w.delay_offsets.mean()
    d = res["delays"] - res["delays"].mean()
    h = res["history"]
    return {
        "workload": w.name, "n_voxels": int(w.n_voxels),
        "forward_ms": t_fwd, "adjoint_ms": t_adj, "adjoint_first_call_ms": t_first, "tangent_ms": t_tan,
        "vjp_ms": t_vjp,
        "forward_plus_adjoint_throughput_vs_forward": t_fwd / (t_fwd + t_adj),
        "forward_autotuned_ms": t_fwd_tuned, "autotuned_tiling": {"m_tiles": tune["m_tiles"],
                                                                 "bin_split": tune["bin_split"]},
        "forward_plus_adjoint_throughput_vs_autotuned_forward": t_fwd_tuned / (t_fwd_tuned + t_adj),
        "autofocus": {"iterations": 50, "seconds": res["seconds"], "c_start": 1480.0, "c_final": res["c"],
                      "delay_rms_error_ns_start": float(np.std(truth) * 1e9),
                      "delay_rms_error_ns_final": float(np.std(d - truth) * 1e9),
                      "objective_gain": h[-1]["J"] / h[0]["J"]},
    }


def run_line_scan(steps, warmup, dev):
    """Line-by-line scan: 128 lines over x in [-19, 19] mm, each from its own transmit focused
    at (x_i, 20 mm), 2048 samples, 512 depths over z in [5, 40] mm (one line-scan plan: one
    spectra launch over the 128 transmit events, one beamforming launch)."""
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    w = synth.line_scan(128, 512)
    rf = w.rf_device(dev)
    scan = fa.LineScan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.f_s, w.width, w.pitch, w.c_0,
                       w.f_bounds, 20e-3, index_gpu=dev.index or 0)
    img = torch.empty((128, 512), dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)
    t_frame = _time_ms(lambda: scan.execute(rf, img, stream=stream), steps, warmup, dev)
    scan.plan.set_profiling(True)
    scan.plan.get_profile()
    for _ in range(steps):
        scan.execute(rf, img, stream=stream)
    torch.cuda.synchronize(dev)
    spec_ms, beam_ms, n = scan.plan.get_profile()
    scan.plan.set_profiling(False)
    info = scan.info()
    scan.close()
    triples = float(info["n_voxels"]) * info["n_f"] * w.n_elements
    beam = beam_ms / max(n, 1)
    return {
        "workload": w.name, "n_lines": 128, "n_voxels": int(w.n_voxels), "frame_ms": t_frame,
        "frames_per_s": 1e3 / t_frame, "mvox_per_s": w.n_voxels / t_frame / 1e3,
        "stages_ms": {"spectra": spec_ms / max(n, 1), "beamform": beam},
        "beamform_weights_per_s": triples / (beam / 1e3),
        "note": "beamforming bound by the vector ALUs: one weight (window, phase sine and cosine) per "
                "(voxel, bin, element), used by one transmit row",
    }


def run_batching(steps, n_frames, n_jobs, dev):
    """Batching form: a Doppler ensemble (n_frames x 5 angles, 1536 samples, 256 x 256,
    flow phantom) and a queue of n_jobs independent CPWC jobs (grids and angle counts drawn
    with a fixed seed) executed together as one captured graph over 4 streams, every plan
    autotuned; mixed precision (fp16) and the same batch in FP32.  The graph is timed with
    one launch per plan, with the jobs of each kernel class as one persistent launch over
    their work items, and with the jobs that share grid and band as row stacks (one launch
    over their stacked transmit rows); the fastest form is the batch's figure."""
    import numpy as np
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    from f_number_amd.tuning import TilingCache

    t_setup = time.perf_counter()
    w0 = synth.doppler_frame(0)
    n_vox_e = w0.n_voxels
    rf_e = torch.empty((n_frames, w0.n_tx, w0.n_elements, w0.n_samples), dtype=torch.float32, device=dev)
    for f in range(n_frames):
        wf = synth.doppler_frame(f)
        wf.n_speckle = 5000  # synthesis time only; the beamforming cost does not depend on the scene
        rf_e[f] = wf.rf_device(dev)
    rng = np.random.default_rng(20260003 + 3000)
    grids = [(128, 128), (256, 384), (512, 512), (512, 768)]
    jobs = []
    for j in range(n_jobs):
        nz, nx = grids[rng.integers(len(grids))]
        na = int(rng.choice([1, 3, 11, 31]))
        w = synth.Workload(f"job{j}", "pw", 1536, np.linspace(-15.0, 15.0, na) if na > 1 else np.zeros(1),
                           np.linspace(-19e-3, 19e-3, nx), np.linspace(5e-3, 40e-3, nz), (2.25e6, 6.75e6),
                           n_speckle=5000, seed=20260003 + 4000 + j)
        jobs.append((w, w.rf_device(dev)))
    torch.cuda.synchronize(dev)
    t_synth = time.perf_counter() - t_setup
    kw = dict(window=fa.windows.hann(), F_number=fa.f_numbers.grating.angle_lb(60, 1.5))

    cache = TilingCache()

    def build(precision):
        plans, rfs, outs = [], [], []
        p = fa.PwPlan(w0.positions_x, w0.positions_z, w0.n_samples, w0.n_elements, w0.angles, w0.f_s, w0.width,
                      w0.pitch, w0.c_0, w0.f_bounds, 0.0, precision=precision, n_frames=n_frames, **kw)
        plans.append(p)
        rfs.append(rf_e)
        outs.append(torch.empty((n_frames, w0.positions_x.size, w0.positions_z.size), dtype=torch.complex64,
                                device=dev))
        for w, rf in jobs:
            plans.append(fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width,
                                   w.pitch, w.c_0, w.f_bounds, 0.0, precision=precision, **kw))
            rfs.append(rf)
            outs.append(torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device=dev))
        t0 = time.perf_counter()
        tunes = [cache.tune(pl, rf, o, repetitions=1) for pl, rf, o in zip(plans, rfs, outs)]
        torch.cuda.synchronize(dev)
        return plans, rfs, outs, tunes, time.perf_counter() - t0

    res = {}
    images = {}
    for precision in ("fp16", "fp32"):
        plans, rfs, outs, tunes, t_tune = build(precision)
        batch = fa.Batch(plans, n_streams=4, persistent=False, stacking=False)
        ms_streams = _time_ms(lambda: batch.execute(rfs, outs), steps, 1, dev)
        nodes_streams = batch.info()["graph_nodes"]
        batch.set_persistent(True)  # same-class jobs as one persistent launch over their work items
        ms_pers = _time_ms(lambda: batch.execute(rfs, outs), steps, 1, dev)
        batch.set_persistent(False)
        batch.set_stacking(True)  # jobs sharing grid and band as one launch over their stacked rows
        ms_stack = _time_ms(lambda: batch.execute(rfs, outs), steps, 1, dev)
        stack_info = batch.info()
        forms = {"stacked": ms_stack, "persistent": ms_pers, "per_plan": ms_streams}
        best = min(forms, key=forms.get)
        batch.set_stacking(best == "stacked")
        batch.set_persistent(best == "persistent")
        ms = _time_ms(lambda: batch.execute(rfs, outs), steps, 1, dev)
        ens_ms = _time_ms(lambda: plans[0].execute(rfs[0], outs[0]), max(2, steps // 2), 1, dev)
        images[precision] = [o.clone() for o in outs]
        info = batch.info()
        voxel_frames = n_frames * n_vox_e + sum(w.n_voxels for w, _ in jobs)
        res[precision] = {"batch_ms": ms, "value": voxel_frames / (ms / 1e3) / 1e6, "unit": "Mvoxel*frames/s",
                          "ensemble_alone_ms": ens_ms, "graph_nodes": info["graph_nodes"],
                          "persistent_ms": ms_pers, "per_plan_launches_ms": ms_streams,
                          "row_stacked_ms": ms_stack, "row_stacks": stack_info["row_stacks"],
                          "stacked_plans": stack_info["stacked_plans"], "form": best,
                          "graph_nodes_per_plan_launches": nodes_streams, "persistent": best == "persistent",
                          "autotune_s": t_tune,
                          "tiling_ensemble": {"m_tiles": tunes[0]["m_tiles"], "bin_split": tunes[0]["bin_split"],
                                              "default_ms": tunes[0]["default_ms"], "best_ms": tunes[0]["best_ms"]},
                          "autotune_gain": sum(t["default_ms"] for t in tunes) / sum(t["best_ms"] for t in tunes)}
        if precision == "fp32":
            # the plain FP32 path: every frame and every job as its own plan execution on one
            # stream (default tilings, no frame batching, no graph)
            single = fa.PwPlan(w0.positions_x, w0.positions_z, w0.n_samples, w0.n_elements, w0.angles, w0.f_s,
                               w0.width, w0.pitch, w0.c_0, w0.f_bounds, 0.0, **kw)
            one = torch.empty((w0.positions_x.size, w0.positions_z.size), dtype=torch.complex64, device=dev)
            per = _time_ms(lambda: [single.execute(rf_e[f], one) for f in range(n_frames)], 2, 1, dev)
            res["ensemble_frames_as_rows_speedup_fp32"] = per / ens_ms
            single.close()
            plain = [fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width,
                               w.pitch, w.c_0, w.f_bounds, 0.0, **kw) for w, _ in jobs]
            jobs_ms = _time_ms(lambda: [pl.execute(rf, o) for pl, (_, rf), o in zip(plain, jobs, outs[1:])], 2, 1,
                               dev)
            for pl in plain:
                pl.close()
            res["plain_fp32_ms"] = per + jobs_ms
        batch.close()
        for pl in plans:
            pl.close()
    num = sum(torch.linalg.vector_norm(a - b).item() ** 2 for a, b in zip(images["fp16"], images["fp32"]))
    den = sum(torch.linalg.vector_norm(b).item() ** 2 for b in images["fp32"])
    return {"ensemble": {"frames": n_frames, "angles": int(w0.n_tx), "grid": [int(w0.positions_z.size),
                                                                              int(w0.positions_x.size)],
                         "n_samples": w0.n_samples},
            "queue_jobs": n_jobs, "voxel_frames_per_batch": int(n_frames * n_vox_e + sum(w.n_voxels for w, _ in jobs)),
            "mixed_fp16": res["fp16"], "fp32": res["fp32"],
            "mixed_vs_fp32_throughput": res["fp16"]["value"] / res["fp32"]["value"],
            "plain_fp32": {"ms": res["plain_fp32_ms"],
                           "value": (n_frames * n_vox_e + sum(w.n_voxels for w, _ in jobs)) / (res["plain_fp32_ms"] / 1e3) / 1e6,
                           "how": "every frame and job as its own FP32 plan execution, default tiling, one stream"},
            "mixed_batched_vs_plain_fp32": res["plain_fp32_ms"] / res["fp16"]["batch_ms"],
            "ensemble_frames_as_rows_speedup_fp32": res["ensemble_frames_as_rows_speedup_fp32"],
            "mixed_rel_l2_vs_fp32": (num / den) ** 0.5, "synthesis_s": t_synth}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: --gpus N > 1 is not supported; a frame "
                                   "shards by voxel tiles across GPUs without exchange"}))
        sys.exit(2)
    import torch

    import f_number_amd as fa
    from f_number_amd import synth

    dev = torch.device("cuda:0")
    if args.copy_calibration_mib:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize(dev)
        del src, dst
    w = {"cpwc_full": synth.cpwc_full, "cpwc_small": synth.cpwc_small}[args.workload]()
    arms = {}
    hann_kw = dict(window=fa.windows.hann(), F_number=fa.f_numbers.grating.angle_lb(60, 1.5))
    precs = [p for p in args.precisions.split(",") if p]
    if args.arm in ("both", "hann"):
        arms["hann_angle_lb"] = run_arm(w, hann_kw, args.steps, args.warmup, dev, keep_image=True,
                                        bin_split=args.bin_split)
    if args.arm in ("both", "boxcar"):
        arms["boxcar_F1.5"] = run_arm(w, dict(window=fa.windows.boxcar(), F_number=fa.f_numbers.constant(1.5)),
                                      args.steps, args.warmup, dev)
    head_key = "hann_angle_lb" if "hann_angle_lb" in arms else "boxcar_F1.5"
    head = arms[head_key]
    # mixed-precision contractions of the headline configuration: speed-up of
    # the beamforming kernel and image error relative to the FP32 image
    prec_arms = {}
    ix_col = sample_column(w)
    cols = {}  # centre lateral column of each arm's image, for the FP64 check of the CPU leg
    if head_key == "hann_angle_lb":
        ref_img = head["image"]
        cols["fp32"] = ref_img[ix_col].cpu().numpy().astype(np.complex128)
        for p in precs:
            r = run_arm(w, dict(hann_kw, precision=p), args.steps, args.warmup, dev, keep_image=True)
            err = (torch.linalg.vector_norm(r["image"] - ref_img) / torch.linalg.vector_norm(ref_img)).item()
            prec_arms[p] = {"ms_per_step": r["ms_per_step"], "beamform_ms": r["beamform_ms"],
                            "prep_ms": r["spectra_ms"], "value": w.n_voxels / (r["ms_per_step"] / 1e3) / 1e6,
                            "beamform_speedup_vs_fp32": head["beamform_ms"] / r["beamform_ms"],
                            "rel_l2_vs_fp32": err, "finite": r["finite"],
                            "roofline": precision_roofline(p, head["info"]["useful_flops_per_frame"],
                                                           r["beamform_ms"])}
            cols[p] = r["image"][ix_col].cpu().numpy().astype(np.complex128)
            del r
    value = w.n_voxels / (head["ms_per_step"] / 1e3) / 1e6
    flops = head["info"]["useful_flops_per_frame"]
    achieved = flops / (head["beamform_ms"] / 1e3) / 1e12
    out = {
        "metric": "Mvoxel*frames/s",
        "value": value,
        "unit": "Mvoxel*frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded point + speckle scatterers, 5 MHz pulse, 30 dB SNR)",
        "config": {
            "workload": w.name,
            "arm": "grating.angle_lb(60,1.5) F-number, Hann receive window, normalization on"
            if head_key == "hann_angle_lb" else "constant F=1.5, boxcar",
            "n_elements": w.n_elements, "n_angles": int(w.n_tx), "n_samples": w.n_samples,
            "grid": [int(w.positions_z.size), int(w.positions_x.size)],
            "f_bounds_hz": list(w.f_bounds), "N_dft": head["info"]["N_dft"], "n_f": head["info"]["n_f"],
            "precision": "fp32", "semantics": "cpu",
        },
        "roofline": {
            "kernel": "beamform_kernel",
            "bound": "mfma",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": measured_traffic(w.name),
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "flops_per_launch": flops,
            "flops_dense_per_launch": head["info"]["flops_per_frame"],
            "flops_note": "8 flop per complex multiply-add over the elements whose weight can be nonzero "
                          "(aperture bounds; every element for an untruncated window such as the CPU Hann)",
            "avg_launch_ms": head["beamform_ms"],
        },
        "stages_ms": {"spectra": head["spectra_ms"], "beamform": head["beamform_ms"]},
        "spectra_roofline": spectra_roofline(w, head["info"]["n_f"], head["spectra_ms"]),
        "arms": {k: {"value": w.n_voxels / (v["ms_per_step"] / 1e3) / 1e6, "ms_per_step": v["ms_per_step"],
                     "beamform_ms": v["beamform_ms"], "spectra_ms": v["spectra_ms"], "finite": v["finite"]}
                 for k, v in arms.items()},
        "precision_arms": prec_arms,
    }
    if not args.no_small:
        out["cpwc_11ang_256x256"] = run_small(args.steps, args.warmup, dev)
    if not args.no_wide:
        out["wide_arrays"] = run_wide(max(3, args.steps // 2), 1, dev)
    if not args.no_diff:
        out["differentiation"] = run_diff(max(3, args.steps // 2), 1, dev)
    if not args.no_line_scan:
        out["line_scan"] = run_line_scan(max(3, args.steps), 2, dev)
    if not args.no_batching:
        out["batching"] = run_batching(max(2, args.steps // 3), args.ensemble_frames, args.queue_jobs, dev)
    saft = None
    if not args.no_saft:
        saft, w_sa, cols_sa, rf_sa = run_saft(max(2, args.steps // 2), 1, dev, precs)
        out["saft"] = saft
    if not args.no_cpu_baseline:
        nthreads = host_cpus()[0]
        if head_key == "hann_angle_lb":
            cb, ref_col = cpu_baseline(w, "hann", "angle_lb", (60, 1.5, 10), nthreads)
            out["cpu_baseline"] = cb
            # error of every arm against the FP64 oracle on the sampled column (GPU arms use
            # the MEX output scale 2/N; the oracle the CPU scale 2)
            n_dft = head["info"]["N_dft"]
            errs = {k: rel_l2(v * n_dft, ref_col) for k, v in cols.items()}
            out["fp64_check"] = {"sample": f"lateral column {ix_col} ({w.positions_z.size} voxels)", "rel_l2": errs}
            for k, e in errs.items():
                if k in prec_arms:
                    prec_arms[k]["rel_l2_vs_fp64"] = e
        if saft is not None:
            errs, cb_sa = saft_fp64_check(
                w_sa, rf_sa, dict(window_tx=("tukey", 0.5), window_rx=("tukey", 0.5)), nthreads, cols_sa,
                saft["N_dft"])
            for k, e in errs.items():
                saft["arms"][k]["rel_l2_vs_fp64"] = e
            saft["fp64_check_sample"] = f"lateral column {sample_column(w_sa)} ({w_sa.positions_z.size} voxels)"
            saft["cpu_baseline"] = cb_sa
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
