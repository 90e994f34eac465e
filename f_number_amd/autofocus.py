"""Gradient-based autofocus of the plane-wave beamformer (differentiation form).

Parameters theta = (c_tof, delay_k): the speed of sound of the time of flight and
per-element receive delay offsets (PwPlan.set_tof).  Objective, maximised:

    J(theta) = sum_v W(v) |I(v; theta)|^2

with a non-negative weight map W: uniform (default), or the transmit-angle
coherence factor
    CF(v) = |sum_a I_a(v)|^2 / (N_a sum_a |I_a(v)|^2),
I_a the single-angle images at the current iterate, refreshed every
`refresh` iterations and held fixed in between (a coherence-weighted image
intensity; W is not differentiated).  dJ = sum_v Re(conj(g) dI) with g = 2 W I; dJ/dc comes from
the tangent pass (PwPlan.execute_tangent) and dJ/d delay from the adjoint
contraction (PwPlan.vjp).  Steps: Adam (Kingma & Ba, 2015) on the scaled
parameters c / (1 m/s) and delay / (1 ns), ascending J.
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib


def angle_coherence(w, rf, plan_kwargs: dict, c_tof: float = 0.0, delays=None, c_apod: float | None = None):
    """CF(v) across transmit angles with single-angle plans (torch float32 (N_x, N_z));
    c_apod: the speed of sound of the plans' apodization (default: the workload's c_0)."""
    import torch

    from .plan import PwPlan

    num = None
    den = None
    for a in range(w.n_tx):
        p = PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles[a:a + 1], w.f_s, w.width,
                   w.pitch, w.c_0 if c_apod is None else c_apod, w.f_bounds, **plan_kwargs)
        p.set_tof(c_tof, delays)
        img = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device=rf.device)
        p.execute(rf[a:a + 1].contiguous(), img)
        num = img.clone() if num is None else num + img
        den = img.abs() ** 2 if den is None else den + img.abs() ** 2
        torch.cuda.synchronize(rf.device)
        p.close()
    return (num.abs() ** 2 / (w.n_tx * den.clamp_min(1e-30))).float().contiguous()


def autofocus(plan, rf, n_iter: int = 50, c0: float = 1480.0, delays0=None, weights=None, lr_c: float = 2.0,
              lr_delay_ns: float = 2.0, fit_delays: bool = True, betas=(0.9, 0.99), log=None,
              coherence=None, refresh: int = 5):
    """Maximise J over (c_tof, delays) with Adam; returns dict(c, delays, history, seconds).

    plan: PwPlan (FP32 kernels); rf: device tensor (N_a, N_el, N_t); weights: torch
    float32 (N_x, N_z) or None (uniform).  coherence=(workload, plan_kwargs) switches
    W to the angle coherence factor of the current iterate, refreshed every `refresh`
    iterations."""
    import torch

    dev = rf.device
    n_el = plan.n_elements
    n_vox = plan.info()["n_voxels"]
    I = torch.empty(n_vox, dtype=torch.complex64, device=dev)
    dI = torch.empty_like(I)
    W = weights.reshape(-1).to(dev, torch.float32) if weights is not None else torch.ones(n_vox, device=dev)
    c = float(c0)
    d = np.zeros(n_el) if delays0 is None else np.asarray(delays0, dtype=np.float64).copy()
    # Adam state on x = (c, delay / 1 ns)
    x = np.concatenate([[c], d * 1e9])
    lr = np.concatenate([[lr_c], np.full(n_el, lr_delay_ns if fit_delays else 0.0)])
    m = np.zeros_like(x)
    v = np.zeros_like(x)
    hist = []
    t0 = time.perf_counter()
    for it in range(n_iter):
        c, d = float(x[0]), x[1:] * 1e-9
        if coherence is not None and it % max(1, refresh) == 0:
            W = angle_coherence(coherence[0], rf, coherence[1], c, d, c_apod=plan.c_0).reshape(-1)
        plan.set_tof(c, d)
        plan.execute_tangent(rf, I, dI)
        g = (2.0 * W * I).contiguous()
        vd, vc = plan.vjp(rf, g, dI)
        J = torch.sum(W * I.real ** 2 + W * I.imag ** 2).item()
        grad = np.concatenate([[vc.item()], vd.cpu().numpy() * 1e-9])  # dJ/dc, dJ/d(delay in ns)
        if not np.all(np.isfinite(grad)):
            raise _lib.FNumberError("autofocus:NonFiniteGradient", f"iteration {it}")
        hist.append({"iter": it, "J": J, "c": c, "grad_c": float(grad[0]),
                     "grad_delay_norm": float(np.linalg.norm(grad[1:]))})
        if log:
            log(hist[-1])
        b1, b2 = betas
        m = b1 * m + (1 - b1) * grad
        v = b2 * v + (1 - b2) * grad * grad
        mh = m / (1 - b1 ** (it + 1))
        vh = v / (1 - b2 ** (it + 1))
        sv = np.sqrt(vh)
        x = x + lr * mh / (sv + 1e-12 * sv.max() + 1e-300)
    torch.cuda.synchronize(dev)
    return {"c": float(x[0]), "delays": x[1:] * 1e-9, "history": hist, "seconds": time.perf_counter() - t0}
