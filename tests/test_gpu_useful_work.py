"""The useful work a plan reports for the roofline (fn_plan_info.useful_flops_per_frame):
8 flop per complex multiply-add over the elements whose weight can be nonzero, per
voxel and bin, times the transmit rows (synthetic aperture: times the active
transmitters).  Pinned against the oracle's own weights, extracted per element from
unit impulses on a one-bin band (the extraction of tests/test_oracle_aperture.py),
on grids small enough that the count samples every voxel."""
import numpy as np
import pytest

import test_oracle_aperture as TA

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fa():
    import f_number_amd as fa

    fa.load()
    return fa


def _fa_window(fa, win):
    name, par = win
    return {"boxcar": fa.windows.boxcar(), "hann": fa.windows.hann(), "triangular": fa.windows.triangular(),
            "tukey": fa.windows.tukey(par) if par else None}[name]


@pytest.mark.parametrize("sem", ["cpu", "mex"])
@pytest.mark.parametrize("win", [("boxcar", 0.0), ("tukey", 0.5), ("hann", 0.0)])
@pytest.mark.parametrize("F", [("constant", 1.0), ("constant", 2.5), ("angle_lb",)])
def test_useful_work_counts_the_oracles_nonzero_weights(fa, oracle, sem, win, F):
    f_target = 4.0e6
    Fo = oracle.fnumber("angle_lb", 60, 1.5, 10) if F[0] == "angle_lb" else oracle.fnumber("constant", F[1])
    Fa = fa.f_numbers.grating.angle_lb(60, 1.5) if F[0] == "angle_lb" else fa.f_numbers.constant(F[1])
    a, N, n1 = TA.extract_pw(oracle, sem, win, Fo, "off", f_target)
    expected = 8.0 * np.count_nonzero(a)  # one bin, one transmit row
    _, _, fb = TA.one_bin(oracle, sem, f_target)
    plan = fa.PwPlan(TA.X, TA.Z, TA.N_T, TA.N_EL, [TA.ANGLE], TA.FS, TA.W, TA.P, TA.C0, fb, 0.0,
                     window=_fa_window(fa, win), F_number=Fa, normalization=fa.normalizations.off(),
                     semantics=sem)
    info = plan.info()
    plan.close()
    print(f"case {sem} {win} {F}: useful {info['useful_flops_per_frame']:.0f} expected {expected:.0f} "
          f"dense {info['flops_per_frame']:.0f}", flush=True)
    assert info["n_f"] == 1 and info["N_dft"] == N
    assert info["useful_flops_per_frame"] == pytest.approx(expected, rel=1e-12)
    assert info["useful_flops_per_frame"] <= info["flops_per_frame"]
