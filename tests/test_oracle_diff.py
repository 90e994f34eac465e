"""Pins of the oracle's derivatives (oracle.h or_das_pw_diff) by what the
mathematics fixes, independently of the formulas they implement:

* dI/dc_tof and the vector-Jacobian products against FP64 central finite
  differences of the oracle's own image (step sizes chosen for a truncation
  error far below the tolerance);
* the adjoint against the dot-product identity <A x, y> = <x, A^H y> with
  <u, y> = Re sum conj(u) y, for random x and y;
* linearity of the adjoint, and its zero response to y = 0.
"""
import numpy as np
import pytest

from f_number_amd import synth


@pytest.fixture(scope="module")
def case():
    w = synth.scaled(synth.cpwc_small(), 6, 5, n_angles=3, n_samples=512, n_speckle=300, n_elements=32)
    rf = w.rf_host().astype(np.float64)
    rng = np.random.default_rng(5)
    delays = rng.normal(0, 20e-9, w.n_elements)
    return w, rf, delays


def _args(w, rf):
    return (w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, w.f_bounds, 0.0)


@pytest.mark.parametrize("win,norm,sem", [("hann", "on", "cpu"), ("boxcar", "off", "cpu"), ("tukey", "on", "mex")])
def test_tangent_and_vjp_match_central_differences(oracle, case, win, norm, sem):
    w, rf, delays = case
    F = oracle.fnumber("angle_lb", 60, 1.5, 10)
    kw = dict(window=(win, 0.5 if win == "tukey" else 0.0), F=F, normalization=norm, semantics=sem)
    c = 1500.0
    base = oracle.das_pw_diff(*_args(w, rf), **kw, c_tof=c, delay_offsets=delays, want_dc=True)
    # J = sum |I|^2: g = 2 I
    r = oracle.das_pw_diff(*_args(w, rf), **kw, c_tof=c, delay_offsets=delays, g=2 * base["image"], want_dc=True)

    def img(cc, dd):
        return oracle.das_pw_diff(*_args(w, rf), **kw, c_tof=cc, delay_offsets=dd)["image"]

    h = 1e-3  # truncation ~ (w tau / c)^3 h^2 / 6, about 1e-7 relative here
    fd_img = (img(c + h, delays) - img(c - h, delays)) / (2 * h)
    assert np.linalg.norm(base["dIdc"] - fd_img) <= 1e-6 * np.linalg.norm(fd_img)
    J = lambda cc, dd: np.sum(np.abs(img(cc, dd)) ** 2)
    fd_c = (J(c + h, delays) - J(c - h, delays)) / (2 * h)
    assert abs(r["vjp_c"] - fd_c) <= 1e-6 * abs(fd_c)
    for k in (0, 13, 31):
        e = np.zeros_like(delays)
        e[k] = 1e-12
        fd = (J(c, delays + e) - J(c, delays - e)) / 2e-12
        assert abs(r["vjp_delay"][k] - fd) <= 1e-5 * np.max(np.abs(r["vjp_delay"]))


@pytest.mark.parametrize("sem", ["cpu", "mex"])
def test_adjoint_dot_product_identity(oracle, case, sem):
    w, rf, delays = case
    F = oracle.fnumber("angle_lb", 60, 1.5, 10)
    kw = dict(window=("hann", 0.0), F=F, normalization="on", semantics=sem, c_tof=1510.0, delay_offsets=delays)
    rng = np.random.default_rng(11)
    x = rng.normal(size=rf.shape)
    Ax = oracle.das_pw_diff(*_args(w, x), **kw)["image"]
    y = rng.normal(size=Ax.shape) + 1j * rng.normal(size=Ax.shape)
    AHy = oracle.das_pw_diff(*_args(w, rf), **kw, y=y)["adjoint"]
    lhs = np.real(np.vdot(Ax, y))
    rhs = np.sum(x * AHy)
    assert abs(lhs - rhs) <= 1e-12 * (np.linalg.norm(Ax) * np.linalg.norm(y))


def test_adjoint_is_linear_in_y(oracle, case):
    w, rf, delays = case
    F = oracle.fnumber("constant", 1.5)
    kw = dict(window=("tukey", 0.5), F=F, normalization="on")
    rng = np.random.default_rng(3)
    shape = (w.positions_z.size, w.positions_x.size)
    y1 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    y2 = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    a = lambda y: oracle.das_pw_diff(*_args(w, rf), **kw, y=y)["adjoint"]
    lin = a(2.0 * y1 - 0.5j * y2)
    assert np.allclose(lin, 2.0 * a(y1) + a(-0.5j * y2), rtol=0, atol=1e-10 * np.max(np.abs(lin)))
    assert np.all(a(np.zeros(shape, complex)) == 0)


def test_delays_shift_like_a_time_shift(oracle, case):
    """A common delay d on every element equals a receive-side time shift:
    the image of x delayed by d equals the image with delay offsets -d... here
    checked through the phase: every valid term gains e^{j w_n d}, so for a
    single-bin band the image is multiplied by exactly that factor."""
    w, rf, delays = case
    F = oracle.fnumber("constant", 1.0)
    N, lb, ub = oracle.grid(w.n_samples, w.n_elements, w.pitch, w.positions_z[0], w.f_s, w.c_0, *w.f_bounds)
    f1 = w.f_s * (lb + 10) / N
    band = (f1 - 1.0, f1 + 1.0)
    kw = dict(window=("hann", 0.0), F=F, normalization="on")
    args = (w.positions_x, w.positions_z, rf, w.f_s, w.angles, w.width, w.pitch, w.c_0, band, 0.0)
    d = 7e-9
    a = oracle.das_pw_diff(*args, **kw, c_tof=w.c_0)["image"]
    b = oracle.das_pw_diff(*args, **kw, c_tof=w.c_0, delay_offsets=np.full(w.n_elements, d))["image"]
    assert np.allclose(b, a * np.exp(1j * 2 * np.pi * f1 * d), rtol=1e-9, atol=1e-12 * np.max(np.abs(a)))


def test_voxel_subset_matches_full_grid(oracle):
    """das_pw_diff on a voxel subset: image and dI/dc equal the full grid's at those
    voxels, and the adjoint of a cotangent supported on the subset equals the full
    grid's adjoint of the same cotangent (zero elsewhere)."""
    rng = np.random.default_rng(31)
    n_el, n_t = 12, 192
    x = np.linspace(-1.5e-3, 1.5e-3, 7)
    z = np.linspace(4e-3, 9e-3, 6)
    rf = rng.standard_normal((2, n_el, n_t))
    kw = dict(window=("tukey", 0.5), F=oracle.fnumber("angle_lb", 60, 1.5, 10), normalization="on")
    d = rng.normal(0, 10e-9, n_el)
    ix = np.array([0, 3, 6, 2, 5])
    iz = np.array([1, 0, 5, 3, 2])
    ys = rng.normal(size=ix.size) + 1j * rng.normal(size=ix.size)
    y = np.zeros((z.size, x.size), complex)
    y[iz, ix] = ys
    args = (x, z, rf, 20.832e6, [-0.05, 0.07], 0.27e-3, 0.3048e-3, 1540.0, (2e6, 7e6), 0.0)
    full = oracle.das_pw_diff(*args, **kw, c_tof=1510.0, delay_offsets=d, y=y, want_dc=True)
    sub = oracle.das_pw_diff(*args, **kw, c_tof=1510.0, delay_offsets=d, y=ys, want_dc=True, sel=(ix, iz))
    assert np.allclose(sub["image"], full["image"][iz, ix], rtol=0, atol=1e-12 * np.abs(full["image"]).max())
    assert np.allclose(sub["dIdc"], full["dIdc"][iz, ix], rtol=0, atol=1e-12 * np.abs(full["dIdc"]).max())
    assert np.linalg.norm(sub["adjoint"] - full["adjoint"]) <= 1e-12 * np.linalg.norm(full["adjoint"])
