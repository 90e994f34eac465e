"""Band-limited spectra of the oracle (das_pw.m:383-384): direct DFT with exact
integer twiddle reduction and Bluestein FFT, against a long-double brute force,
numpy's FFT, a pure tone and Parseval."""
import numpy as np
import pytest


def brute(x, N, n_lb, n_ub):
    x = np.asarray(x, dtype=np.longdouble)
    t = np.arange(x.shape[-1], dtype=np.longdouble)
    out = []
    for n in range(n_lb, n_ub + 1):
        ang = 2 * np.pi * np.longdouble(n) * t / N
        out.append((x * np.cos(ang)).sum(-1) - 1j * (x * np.sin(ang)).sum(-1))
    return np.array(out, dtype=np.complex128).T


@pytest.mark.parametrize("n_t,N", [(50, 61), (97, 127), (200, 211)])
@pytest.mark.parametrize("mode", [0, 1])
def test_against_brute_force(oracle, n_t, N, mode):
    rng = np.random.default_rng(n_t)
    x = rng.standard_normal((3, n_t))
    got = oracle.spectra(x, N, 3, N // 2, mode=mode)
    ref = brute(x, N, 3, N // 2)
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
    assert np.allclose(got, np.fft.fft(x, N, axis=1)[:, 3:N // 2 + 1], rtol=1e-11, atol=1e-10)


def test_pure_tone_and_parseval(oracle):
    N, m0 = 301, 37
    t = np.arange(N)
    x = np.cos(2 * np.pi * m0 * t / N)[None]
    X = oracle.spectra(x, N, 1, N // 2, mode=1)[0]
    assert abs(X[m0 - 1] - N / 2) < 1e-9 and np.max(np.abs(np.delete(X, m0 - 1))) < 1e-9
    rng = np.random.default_rng(1)
    y = rng.standard_normal((1, N))
    Y = oracle.spectra(y, N, 1, N // 2, mode=0)[0]
    Y0 = y.sum()
    assert np.isclose((y ** 2).sum(), (Y0 ** 2 + 2 * np.sum(np.abs(Y) ** 2)) / N, rtol=1e-12)
