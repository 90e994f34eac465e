"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from
the FP64 oracle on small seeded inputs).

CPU: the oracle still reproduces every fixture, and the product's host code (the
C++ F-number port and window samples behind fn_f_number_values /
fn_window_samples) equals the F-number and window tables bit for bit.
GPU: the HIP path reproduces the image, dI/dc, adjoint and vector-Jacobian
fixtures at the plain-FP32 bar (rel-L2 <= 1e-4, peak error <= -80 dB; 1e-4 for
the differentiation quantities).
"""
import os
import sys

import numpy as np
import pytest

from conftest import peak_db, rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLDEN)

import make_golden as MG  # noqa: E402


def load(name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


@pytest.fixture(scope="module")
def regenerated(oracle):
    return MG.compute()


@pytest.mark.parametrize("name", ["fnumbers", "windows", "das_pw", "das_sa", "diff"])
def test_oracle_reproduces_fixture(regenerated, name):
    fix = load(name)
    new = regenerated[name]
    assert set(fix) == set(new)
    for k, v in fix.items():
        got = np.asarray(new[k])
        assert got.shape == v.shape, k
        if np.issubdtype(v.dtype, np.complexfloating) or np.issubdtype(v.dtype, np.floating):
            scale = max(float(np.max(np.abs(v))), 1e-300)
            assert float(np.max(np.abs(got - v))) <= 1e-12 * scale, k
        else:
            assert np.array_equal(got, v), k


PROD_F = {
    "constant_1.5": lambda fa: fa.f_numbers.constant(1.5),
    "angle_lb_60_1.5_10": lambda fa: fa.f_numbers.grating.angle_lb(60, 1.5, 10),
    "angle_lb_70_1.5_10": lambda fa: fa.f_numbers.grating.angle_lb(70, 1.5, 10),
    "distance_constant_10_2": lambda fa: fa.f_numbers.distance.constant(10, 2.0),
    "szabo_0.918": lambda fa: fa.f_numbers.directivity.szabo(0.918),
    "directivity_soft_0.918_6": lambda fa: fa.f_numbers.directivity.soft(0.918, 6),
    "directivity_rigid_0.918_6": lambda fa: fa.f_numbers.directivity.rigid(0.918, 6),
    "perrot_0.918": lambda fa: fa.f_numbers.directivity.perrot(0.918),
    "distance_none": lambda fa: fa.f_numbers.distance.none(),
}


@pytest.mark.parametrize("key", sorted(PROD_F))
def test_product_fnumbers_match_fixture(key):
    import f_number_amd as fa

    fix = load("fnumbers")
    got = PROD_F[key](fa).compute_values(fix["r"])
    assert np.array_equal(got, fix[key])


def test_product_windows_match_fixture():
    import f_number_amd as fa

    fix = load("windows")
    make = {"boxcar": lambda p: fa.windows.boxcar(), "hann": lambda p: fa.windows.hann(),
            "tukey": lambda p: fa.windows.tukey(p), "triangular": lambda p: fa.windows.triangular()}
    n = 0
    for key, ref in fix.items():
        if key == "t":
            continue
        name, p, sem = key.split("_")
        got = make[name](float(p)).compute_samples(fix["t"], semantics=sem)
        assert np.array_equal(got, ref), key
        n += 1
    assert n == 10


# ---------------------------------------------------------------- GPU
TOL_L2, TOL_DB = 1e-4, -80.0


def _fa():
    import f_number_amd as fa

    fa.load()
    return fa


def _F(fa, spec):
    if spec[0] == "angle_lb":
        return fa.f_numbers.grating.angle_lb(*spec[1:])
    return fa.f_numbers.constant(spec[1])


def _W(fa, spec):
    name, p = spec
    return {"boxcar": lambda: fa.windows.boxcar(), "hann": lambda: fa.windows.hann(),
            "tukey": lambda: fa.windows.tukey(p), "triangular": lambda: fa.windows.triangular()}[name]()


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(MG.PW_CASES))
def test_gpu_das_pw_matches_fixture(case):
    fa = _fa()
    fix = load("das_pw")
    win, F, norm, sem = MG.PW_CASES[case]
    G = MG.GEO
    print(f"case golden das_pw {case}", flush=True)
    img = fa.das_pw(fix["x"], fix["z"], np.transpose(fix["rf"], (2, 1, 0)), G["f_s"], fix["angles"], G["width"],
                    G["pitch"], G["c_0"], G["f_bounds"], 0, _W(fa, win), _F(fa, F),
                    fa.normalizations.on() if norm == "on" else fa.normalizations.off(), fa.platforms.gpu(0),
                    semantics=sem, output_scale="cpu")
    e, db = rel_l2(img, fix[case]), peak_db(img, fix[case])
    print(f"  rel_l2 {e:.2e} peak {db:.1f} dB", flush=True)
    assert e <= TOL_L2 and db <= TOL_DB


@pytest.mark.gpu
def test_gpu_das_sa_matches_fixture():
    fa = _fa()
    fix = load("das_sa")
    G = MG.GEO
    print("case golden das_sa", flush=True)
    img = fa.das_sa(fix["x"], fix["z"], np.transpose(fix["rf"], (2, 1, 0)), G["f_s"], G["width"], G["pitch"],
                    G["c_0"], G["f_bounds"], 0, fa.windows.tukey(0.5), fa.windows.tukey(0.5),
                    fa.f_numbers.grating.angle_lb(60, 1.5, 10), fa.f_numbers.grating.angle_lb(60, 1.5, 10),
                    fa.normalizations.on(), fa.normalizations.on(), fa.platforms.gpu(0), output_scale="cpu")
    img = img[0] if isinstance(img, tuple) else img
    e, db = rel_l2(img, fix["tukey_anglelb_on_cpu"]), peak_db(img, fix["tukey_anglelb_on_cpu"])
    print(f"  rel_l2 {e:.2e} peak {db:.1f} dB", flush=True)
    assert e <= TOL_L2 and db <= TOL_DB


@pytest.mark.gpu
def test_gpu_differentiation_matches_fixture():
    import torch

    fa = _fa()
    fix = load("das_pw")
    dif = load("diff")
    G = MG.GEO
    print("case golden differentiation", flush=True)
    plan = fa.PwPlan(fix["x"], fix["z"], fix["rf"].shape[2], fix["rf"].shape[1], fix["angles"], G["f_s"],
                     G["width"], G["pitch"], G["c_0"], G["f_bounds"], 0.0, window=fa.windows.tukey(0.5),
                     F_number=fa.f_numbers.grating.angle_lb(60, 1.5, 10), normalization=fa.normalizations.on(),
                     semantics="cpu", output_scale="cpu")
    plan.set_tof(float(dif["c_tof"]), dif["delays"])
    rf = torch.from_numpy(fix["rf"]).cuda()
    n = (fix["x"].size, fix["z"].size)
    img = torch.empty(n, dtype=torch.complex64, device="cuda")
    dimg = torch.empty_like(img)
    plan.execute_tangent(rf, img, dimg)
    y = torch.from_numpy(np.ascontiguousarray(dif["y"].T).astype(np.complex64)).cuda()
    adj = plan.adjoint(y)
    vd, vc = plan.vjp(rf, y, dimg)
    torch.cuda.synchronize()
    errs = {
        "image": rel_l2(img.cpu().numpy().T, dif["image"]),
        "dIdc": rel_l2(dimg.cpu().numpy().T, dif["dIdc"]),
        "adjoint": rel_l2(adj.cpu().numpy(), dif["adjoint"]),
        "vjp_delay": rel_l2(vd.cpu().numpy(), dif["vjp_delay"]),
        "vjp_c": abs(float(vc.item()) - float(dif["vjp_c"])) / abs(float(dif["vjp_c"])),
    }
    plan.close()
    print("  " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert all(v <= 1e-4 for v in errs.values()), errs
