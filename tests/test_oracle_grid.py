"""Frequency grid (das_pw.m:343-358; gpu_bf_das_pw_rf.cu:525-552): integer closed forms."""
import math

import numpy as np
import pytest

P, C0, FS = 0.3048e-3, 1540.0, 20.832e6


@pytest.mark.parametrize("n_t,band,N,lb,ub", [
    (2048, (2.25e6, 6.75e6), 2509, 271, 812),   # CPWC 11 angles
    (3328, (1e6, 9e6), 3789, 182, 1636),        # CPWC 75 angles
    (3000, (2.25e6, 6.75e6), 3461, 374, 1121),  # SAFT
    (1536, (2.25e6, 6.75e6), 1997, 216, 647),   # Doppler ensemble
])
def test_grid_at_the_workloads(oracle, n_t, band, N, lb, ub):
    assert oracle.grid(n_t, 128, P, 5e-3, FS, C0, *band) == (N, lb, ub)
    # independent restatement: pad covers the largest extra one-way path at z0
    span = 127 * P
    pad = math.ceil((math.hypot(span, 5e-3) - 5e-3) * FS / C0)
    assert N % 2 == 1 and N - (n_t + pad) in (0, 1)
    assert lb == math.ceil(band[0] * N / FS) and ub == math.floor(band[1] * N / FS)


def test_mex_pads_with_second_depth(oracle):
    """quirk: the MEX pads with positions_z[1] (5.137 mm here) -> N_dft 2507 instead of 2509."""
    assert oracle.grid(2048, 128, P, 5e-3 + 35e-3 / 255, FS, C0, 2.25e6, 6.75e6)[0] == 2507


def test_product_grid_matches_oracle(oracle):
    import ctypes as C

    from f_number_amd import _lib
    from f_number_amd.das import pw_params
    import f_number_amd as fa

    rng = np.random.default_rng(3)
    for _ in range(25):
        n_t = int(rng.integers(64, 5000))
        z = np.sort(rng.uniform(1e-3, 50e-3, 4))
        fb = (float(rng.uniform(0.2e6, 3e6)), float(rng.uniform(4e6, 10e6)))
        for sem, zp in (("cpu", z[0]), ("mex", z[1])):
            P_, keep, info = pw_params(np.zeros(3), z, n_t, 128, [0.0], FS, 0.27e-3, P, C0, fb, 0,
                                       fa.windows.boxcar(), fa.f_numbers.constant(1.0), fa.normalizations.on(),
                                       semantics=sem)
            assert (info["N_dft"], info["n_lb"], info["n_ub"]) == oracle.grid(n_t, 128, P, zp, FS, C0, *fb)
