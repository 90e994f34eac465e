"""GPU parity of the differentiation form (fnumber_hip.h fn_plan_set_tof,
fn_plan_execute_tangent, fn_plan_adjoint, fn_plan_vjp) against the FP64 oracle
(or_das_pw_diff, itself pinned by central differences and the dot-product
identity in test_oracle_diff.py).

Bars, FP32 kernels vs FP64:
  image with c_tof != c_0 and delay offsets        rel-L2 <= 1e-4 (the FP32 bar)
  tangent dI/dc_tof                                 rel-L2 <= 1e-4
  adjoint A^H y                                     rel-L2 <= 1e-4
  dot product <A x, y> vs <x, A^H y> (GPU both)     relative 1e-6 (y correlated with A x)
  vjp (delays, c_tof)                               relative 1e-4 of the largest entry
"""
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


def _case(n_el, n_ang, seed):
    from f_number_amd import synth

    w = synth.scaled(synth.cpwc_small(), 14, 12, n_angles=min(n_ang, 11), n_samples=1024, n_speckle=1500,
                     n_elements=n_el)
    if n_ang > 11:
        w.angles_deg = np.linspace(-10, 10, n_ang)
    w.seed = seed
    rng = np.random.default_rng(seed)
    delays = rng.normal(0, 20e-9, n_el)
    return w, w.rf_host(), delays


_CASES = {}


def case(n_el, n_ang):
    key = (n_el, n_ang)
    if key not in _CASES:
        _CASES[key] = _case(n_el, n_ang, 7000 + n_el + n_ang)
    return _CASES[key]


def _win(fa, name):
    return {"boxcar": fa.windows.boxcar(), "hann": fa.windows.hann(), "tukey": fa.windows.tukey(0.5),
            "triangular": fa.windows.triangular()}[name]


def _plan(fa, w, win, sem, norm="on"):
    return fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=_win(fa, win), F_number=fa.f_numbers.grating.angle_lb(60, 1.5),
                     normalization=fa.normalizations.on() if norm == "on" else fa.normalizations.off(),
                     semantics=sem, output_scale="cpu")


def _okw(oracle, win, sem, norm="on"):
    return dict(window=(win, 0.5 if win == "tukey" else 0.0), F=oracle.fnumber("angle_lb", 60, 1.5, 10),
                normalization=norm, semantics=sem)


def _img_np(t):
    """device complex64 (N_x, N_z) -> numpy [N_z, N_x]"""
    return t.cpu().numpy().T


def _img_dev(torch, a):
    """numpy [N_z, N_x] -> device complex64 (N_x, N_z)"""
    return torch.from_numpy(np.ascontiguousarray(a.T).astype(np.complex64)).cuda()


# (128, tukey, cpu): every element in support on a full 128-element array (the adjoint's FULL variant)
CFG = [(64, 3, "hann", "cpu", "on"), (128, 11, "tukey", "mex", "on"), (32, 40, "boxcar", "cpu", "off"),
       (100, 20, "triangular", "cpu", "off"), (128, 5, "tukey", "cpu", "on")]


@pytest.mark.parametrize("n_el,n_ang,win,sem,norm", CFG)
def test_tangent_matches_oracle(fa, oracle, torch, n_el, n_ang, win, sem, norm):
    w, rf, delays = case(n_el, n_ang)
    c = 1512.0
    print(f"case tangent n_el={n_el} n_ang={n_ang} win={win} sem={sem} norm={norm}", flush=True)
    plan = _plan(fa, w, win, sem, norm)
    plan.set_tof(c, delays)
    rfd = torch.from_numpy(rf).cuda()
    n = (w.positions_x.size, w.positions_z.size)
    img = torch.empty(n, dtype=torch.complex64, device="cuda")
    dimg = torch.empty_like(img)
    plain = torch.empty_like(img)
    plan.execute_tangent(rfd, img, dimg)
    plan.execute(rfd, plain)
    torch.cuda.synchronize()
    ref = oracle.das_pw_diff(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                             w.f_bounds, 0.0, **_okw(oracle, win, sem, norm), c_tof=c, delay_offsets=delays,
                             want_dc=True)
    e_img, e_pl, e_d = rel_l2(_img_np(img), ref["image"]), rel_l2(_img_np(plain), ref["image"]), \
        rel_l2(_img_np(dimg), ref["dIdc"])
    print(f"  image {e_img:.2e} (plain execute {e_pl:.2e}) dI/dc {e_d:.2e}", flush=True)
    plan.close()
    assert e_img <= 1e-4 and e_pl <= 1e-4 and e_d <= 1e-4


@pytest.mark.parametrize("n_el,n_ang,win,sem,norm", CFG)
def test_adjoint_matches_oracle(fa, oracle, torch, n_el, n_ang, win, sem, norm):
    w, rf, delays = case(n_el, n_ang)
    c = 1495.0
    print(f"case adjoint n_el={n_el} n_ang={n_ang} win={win} sem={sem} norm={norm}", flush=True)
    plan = _plan(fa, w, win, sem, norm)
    plan.set_tof(c, delays)
    rng = np.random.default_rng(n_el + n_ang)
    y = rng.normal(size=(w.positions_z.size, w.positions_x.size)) + 1j * rng.normal(
        size=(w.positions_z.size, w.positions_x.size))
    got = plan.adjoint(_img_dev(torch, y)).cpu().numpy()
    ref = oracle.das_pw_diff(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                             w.f_bounds, 0.0, **_okw(oracle, win, sem, norm), c_tof=c, delay_offsets=delays,
                             y=y)["adjoint"]
    e = rel_l2(got, ref)
    print(f"  adjoint rel_l2 {e:.2e}", flush=True)
    plan.close()
    assert e <= 1e-4


@pytest.mark.parametrize("n_el,n_ang", [(128, 11), (64, 20)])
def test_dot_product_identity_on_gpu(fa, torch, n_el, n_ang):
    """<A x, y> = <x, A^H y> with both sides from the GPU kernels (sums in FP64);
    y = A x + 0.1 noise keeps the inner product well conditioned."""
    w, rf, delays = case(n_el, n_ang)
    plan = _plan(fa, w, "hann", "cpu")
    plan.set_tof(1530.0, delays)
    x = torch.from_numpy(rf).cuda()
    Ax = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    plan.execute(x, Ax)
    torch.cuda.synchronize()
    gen = torch.Generator(device="cuda").manual_seed(5)
    noise = torch.randn(Ax.shape, dtype=torch.complex64, device="cuda", generator=gen)
    y = (Ax + 0.1 * noise * Ax.abs().mean()).contiguous()
    AHy = plan.adjoint(y)
    torch.cuda.synchronize()
    lhs = torch.sum((Ax.conj() * y).real.double()).item()
    rhs = torch.sum(x.double() * AHy.double()).item()
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"case dot n_el={n_el} n_ang={n_ang}: <Ax,y>={lhs:.9e} <x,AHy>={rhs:.9e} rel={rel:.2e}", flush=True)
    plan.close()
    assert rel <= 1e-6


@pytest.mark.parametrize("n_el,n_ang,win,sem,norm", CFG[:3])
def test_vjp_matches_oracle(fa, oracle, torch, n_el, n_ang, win, sem, norm):
    w, rf, delays = case(n_el, n_ang)
    c = 1488.0
    print(f"case vjp n_el={n_el} n_ang={n_ang} win={win} sem={sem} norm={norm}", flush=True)
    plan = _plan(fa, w, win, sem, norm)
    plan.set_tof(c, delays)
    rfd = torch.from_numpy(rf).cuda()
    n = (w.positions_x.size, w.positions_z.size)
    img = torch.empty(n, dtype=torch.complex64, device="cuda")
    dimg = torch.empty_like(img)
    plan.execute_tangent(rfd, img, dimg)
    g = (2 * img).contiguous()  # J = sum |I|^2
    vd, vc = plan.vjp(rfd, g, dimg)
    torch.cuda.synchronize()
    ref = oracle.das_pw_diff(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                             w.f_bounds, 0.0, **_okw(oracle, win, sem, norm), c_tof=c, delay_offsets=delays,
                             g=2 * ref_image(oracle, w, rf, win, sem, norm, c, delays), want_dc=True)
    vd = vd.cpu().numpy()
    e_d = np.max(np.abs(vd - ref["vjp_delay"])) / np.max(np.abs(ref["vjp_delay"]))
    e_c = abs(vc.item() - ref["vjp_c"]) / abs(ref["vjp_c"])
    print(f"  dJ/d delay max rel {e_d:.2e}  dJ/dc rel {e_c:.2e} ({vc.item():.6e} vs {ref['vjp_c']:.6e})", flush=True)
    plan.close()
    assert e_d <= 1e-4 and e_c <= 1e-4


def ref_image(oracle, w, rf, win, sem, norm, c, delays):
    return oracle.das_pw_diff(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                              w.f_bounds, 0.0, **_okw(oracle, win, sem, norm), c_tof=c,
                              delay_offsets=delays)["image"]


def test_set_tof_default_restores_plain_image(fa, torch):
    """set_tof(0, None) returns to the plan's c_0 without offsets.  Bins split
    across blocks are summed with atomics, so images agree to rounding, not bitwise."""
    w, rf, delays = case(64, 3)
    rfd = torch.from_numpy(rf).cuda()
    a = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    b = torch.empty_like(a)
    c = torch.empty_like(a)
    fresh = _plan(fa, w, "hann", "cpu")
    fresh.execute(rfd, a)
    plan = _plan(fa, w, "hann", "cpu")
    plan.set_tof(1400.0, delays)
    plan.execute(rfd, c)
    plan.set_tof(0.0, None)
    plan.execute(rfd, b)
    torch.cuda.synchronize()
    fresh.close()
    plan.close()
    assert (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(a)).item() < 1e-6
    assert (torch.linalg.vector_norm(a - c) / torch.linalg.vector_norm(a)).item() > 1e-2


def test_virtual_source_tangent_and_adjoint_match_oracle(fa, oracle, torch):
    """The differentiation form with diverging / focused transmits (transmit time
    (|r - r_vs| -+ |r_vs|) / c scales as 1 / c like the plane wave's)."""
    w, rf, delays = case(64, 3)
    vs = np.array([[-3e-3, -10e-3], [1e-3, -20e-3], [0.0, 15e-3]])
    c = 1508.0
    print("case virtual-source tangent/adjoint", flush=True)
    plan = fa.PwPlan(w.positions_x, w.positions_z, w.n_samples, w.n_elements, w.angles, w.f_s, w.width, w.pitch,
                     w.c_0, w.f_bounds, 0.0, window=fa.windows.tukey(0.5),
                     F_number=fa.f_numbers.grating.angle_lb(60, 1.5), semantics="cpu", output_scale="cpu",
                     virtual_sources=vs)
    plan.set_tof(c, delays)
    rfd = torch.from_numpy(rf).cuda()
    n = (w.positions_x.size, w.positions_z.size)
    img = torch.empty(n, dtype=torch.complex64, device="cuda")
    dimg = torch.empty_like(img)
    plan.execute_tangent(rfd, img, dimg)
    rng = np.random.default_rng(3)
    y = rng.normal(size=(w.positions_z.size, w.positions_x.size)) + 1j * rng.normal(size=(w.positions_z.size,
                                                                                          w.positions_x.size))
    adj = plan.adjoint(_img_dev(torch, y)).cpu().numpy()
    torch.cuda.synchronize()
    ref = oracle.das_pw_diff(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds,
                             0.0, **_okw(oracle, "tukey", "cpu", "on"), c_tof=c, delay_offsets=delays, y=y,
                             want_dc=True, virtual_sources=vs)
    e = (rel_l2(_img_np(img), ref["image"]), rel_l2(_img_np(dimg), ref["dIdc"]), rel_l2(adj, ref["adjoint"]))
    print(f"  image {e[0]:.2e} dI/dc {e[1]:.2e} adjoint {e[2]:.2e}", flush=True)
    plan.close()
    assert max(e) <= 1e-4


def test_adjoint_ragged_samples(fa, oracle, torch):
    """Sample count not a multiple of 4 (the inverse DFT's scalar store path) and a
    channel count not a multiple of 64 (a partial channel tile): adjoint against the
    FP64 oracle and the dot-product identity."""
    from f_number_amd import synth

    n_el, n_ang, n_t = 48, 7, 1021
    w = synth.scaled(synth.cpwc_small(), 13, 11, n_angles=n_ang, n_samples=n_t, n_speckle=800, n_elements=n_el)
    w.seed = 7700
    rf = w.rf_host()
    delays = np.random.default_rng(7701).normal(0, 20e-9, n_el)
    print(f"case adjoint ragged n_el={n_el} n_ang={n_ang} n_t={n_t}", flush=True)
    plan = _plan(fa, w, "hann", "cpu")
    plan.set_tof(1505.0, delays)
    rng = np.random.default_rng(7702)
    y = rng.normal(size=(w.positions_z.size, w.positions_x.size)) + 1j * rng.normal(
        size=(w.positions_z.size, w.positions_x.size))
    got = plan.adjoint(_img_dev(torch, y)).cpu().numpy()
    assert got.shape == (n_ang, n_el, n_t)
    ref = oracle.das_pw_diff(w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0,
                             w.f_bounds, 0.0, **_okw(oracle, "hann", "cpu"), c_tof=1505.0, delay_offsets=delays,
                             y=y)["adjoint"]
    e = rel_l2(got, ref)
    # dot-product identity with y = A x + 0.1 noise (a well-conditioned inner product)
    x = torch.from_numpy(rf).cuda()
    Ax = torch.empty((w.positions_x.size, w.positions_z.size), dtype=torch.complex64, device="cuda")
    plan.execute(x, Ax)
    torch.cuda.synchronize()
    gen = torch.Generator(device="cuda").manual_seed(7703)
    noise = torch.randn(Ax.shape, dtype=torch.complex64, device="cuda", generator=gen)
    yd = (Ax + 0.1 * noise * Ax.abs().mean()).contiguous()
    AHy = plan.adjoint(yd)
    torch.cuda.synchronize()
    lhs = torch.sum((Ax.conj() * yd).real.double()).item()
    rhs = torch.sum(x.double() * AHy.double()).item()
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"  adjoint rel_l2 {e:.2e}, dot identity rel {rel:.2e}", flush=True)
    plan.close()
    assert e <= 1e-4 and rel <= 1e-6
