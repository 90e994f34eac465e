"""Mathematical pins of the oracle's aperture, apodization and normalization stage
(das_pw.m:387, 399-411, 436-457; das_sa.m:384-414, 438-480; MEX semantics
das_kernel_phase_shifts.cu:90-125 with gpu_bf_windows.cu:15-24).

Weight extraction.  When channel k carries a unit impulse at t = 0 and nothing
else, its spectrum is X_k(n) = 1 at every bin, so on a one-bin band the image is
    I(v) = 2 a_k(v) e^{j w tau_k(v)}.
The same acquisition imaged with unit weights (F = 0, boxcar, normalization off:
every element weight 1, checked here as |I_unit| = 2) is 2 e^{j w tau_k(v)}, so
    a_k(v) = I(v) / I_unit(v),
a real number, for every voxel of the grid at once.

Expected weights are built by scanning the element centres, without the index
arithmetic of the reference (ceil / floor of M + (x -+ h) / pitch):
  - membership: |x_k - x| <= h with h = z / (2 F) (F = 0: every element);
  - an empty member set is an empty aperture: the bin is skipped (image 0);
  - W_L / W_R: distance from x to the outer face of the leftmost / rightmost member;
  - side: W_L for x_k < x, W_R otherwise;
  - CPU: boxcar is [|x_k - x| < h]; Hann and triangular are not truncated;
    MEX: every window is 0 where |x_k - x| >= W of its side;
  - normalization on: a_k / sum_j a_j.
The window shapes are written as cos^2 forms here (the oracle evaluates
(1 + cos) / 2), so a shape slip would also show.

test_planted_mistakes_are_caught builds the oracle with a planted mistake
(ceil and floor exchanged, W_L and W_R exchanged) and checks that these pins fail.
"""
import math

import numpy as np
import pytest

FS, C0, P, W = 20.832e6, 1540.0, 0.3048e-3, 0.27e-3
N_EL, N_T = 16, 64
XK = (np.arange(N_EL) - (N_EL - 1) / 2) * P

# lateral positions: element centres, element faces, the midpoint of the two centre
# elements, outside the array on both sides, generic positions; and their mirrors
_X0 = np.array([XK[0], XK[3], XK[7], XK[8], XK[15],
                XK[0] - W / 2, XK[3] + W / 2, XK[8] - W / 2, XK[15] + W / 2, XK[12] - W / 2,
                0.0, XK[4] + P / 2, -3.5e-3, 3.1e-3, 0.4111e-3, -1.0007e-3])
X = np.concatenate([_X0, -_X0[::-1]])
Z = np.array([2.0e-3, 3.3e-3, 5.1e-3, 8.7e-3])
ANGLE = 0.13  # any steering angle: it only enters the phase, which the ratio removes


def _shape(win, t, truncated):
    """Window value at |t| (closed forms: +windows/*.m compute_samples_scalar,
    gpu_bf_windows.cu:30-103)."""
    name, par = win
    if name == "boxcar":
        return 1.0 if t < 1 else 0.0
    if name == "tukey":
        if t >= 1:
            return 0.0
        if t <= 1 - par:
            return 1.0
        return math.cos(math.pi * (t - 1 + par) / (2 * par)) ** 2
    if truncated and t >= 1:
        return 0.0
    if name == "hann":
        return math.cos(math.pi * t / 2) ** 2
    if name == "triangular":
        return 1.0 - t
    raise ValueError(name)


def expected_weights(x, z, F, win, sem, norm):
    """Weights of one voxel from an element scan; None for an empty aperture."""
    h = math.inf if F == 0 else z / (2 * F)
    d = XK - x
    members = np.flatnonzero(np.abs(d) <= h)
    if members.size == 0:
        return None
    WL = x - (XK[members[0]] - W / 2)   # outer face of the leftmost member
    WR = (XK[members[-1]] + W / 2) - x  # outer face of the rightmost member
    a = np.empty(N_EL)
    for k in range(N_EL):
        if sem == "cpu" and win[0] == "boxcar":
            a[k] = 1.0 if abs(d[k]) < h else 0.0
            continue
        Ws = WL if d[k] < 0 else WR
        if sem == "mex":
            a[k] = 0.0 if abs(d[k]) >= Ws else _shape(win, abs(d[k]) / Ws, True)
        else:
            assert Ws != 0.0, "degenerate half-width: the parameters must avoid it"
            a[k] = _shape(win, abs(d[k] / Ws), False)
    if norm == "on":
        a = a / a.sum()
    return a


def one_bin(O, sem, f_target):
    """The one-bin band around f_target for the semantics' DFT order
    (N_pad from positions_z(1), or positions_z[1] in MEX semantics)."""
    z_pad = Z[1] if sem == "mex" else Z[0]
    N = O.grid(N_T, N_EL, P, z_pad, FS, C0, 1e6, 9e6)[0]
    n1 = round(f_target * N / FS)
    return N, n1, ((n1 - 0.4) * FS / N, (n1 + 0.4) * FS / N)


_UNIT = {}


def _impulse(k):
    rf = np.zeros((1, N_EL, N_T))
    rf[0, k, 0] = 1.0
    return rf


def unit_image(O, sem, f_target, k):
    key = (id(O.lib()), sem, f_target, k)
    if key not in _UNIT:
        N, n1, fb = one_bin(O, sem, f_target)
        img, info = O.das_pw(X, Z, _impulse(k), FS, [ANGLE], W, P, C0, fb, 0, ("boxcar", 0.0),
                             O.fnumber("constant", 0.0), "off", sem)
        assert info["n_lb"] == info["n_ub"] == n1
        _UNIT[key] = img
    return _UNIT[key]


def extract_pw(O, sem, win, F, norm, f_target):
    """a[k, iz, ix] of das_pw for every element k, from unit impulses."""
    N, n1, fb = one_bin(O, sem, f_target)
    a = np.empty((N_EL, Z.size, X.size))
    for k in range(N_EL):
        img, _ = O.das_pw(X, Z, _impulse(k), FS, [ANGLE], W, P, C0, fb, 0, win, F, norm, sem)
        ref = unit_image(O, sem, f_target, k)
        assert np.max(np.abs(np.abs(ref) - 2.0)) <= 1e-13  # unit weights
        q = img / ref
        assert np.max(np.abs(q.imag)) <= 1e-12 * max(1.0, np.max(np.abs(q.real)))
        a[k] = q.real
    return a, N, n1


def f_value(O, F, f):
    return float(O.fnumber_values(F, [P * f / C0])[0])


WINDOWS = [("boxcar", 0.0), ("tukey", 0.5), ("tukey", 0.25), ("hann", 0.0), ("triangular", 0.0)]
# (F-number, frequency of the one-bin band): constant F at 0 (full aperture), 1, 3, 7;
# grating.angle_lb(60, 1.5, 10) on its rising branch (3 MHz) and at F_ub (5 MHz)
F_CASES = [("constant", 0.0, 5e6), ("constant", 1.0, 5e6), ("constant", 3.0, 5e6), ("constant", 7.0, 5e6),
           ("angle_lb", None, 3e6), ("angle_lb", None, 5e6)]


def _fobj(O, fc):
    return O.fnumber("constant", fc[1]) if fc[0] == "constant" else O.fnumber("angle_lb", 60, 1.5, 10)


def check_pw_case(O, sem, win, fc, norm):
    """Max |extracted - expected| over the grid (np.inf on a structural mismatch)."""
    F = _fobj(O, fc)
    a, N, n1 = extract_pw(O, sem, win, F, norm, fc[2])
    f = FS * n1 / N
    Fn = f_value(O, F, f)
    err = 0.0
    for iz, z in enumerate(Z):
        # das_pw.m:374-377: the depth clamp of the CPU path (absent from the MEX)
        Fz = min(Fn, math.sqrt(f * z / (2.88 * C0))) if sem == "cpu" else Fn
        for ix, x in enumerate(X):
            e = expected_weights(x, z, Fz, win, sem, norm)
            got = a[:, iz, ix]
            if e is None:
                if np.any(got != 0.0):
                    return np.inf  # an empty aperture skips the bin (das_pw.m:403,414)
                continue
            err = max(err, float(np.max(np.abs(got - e))))
    return err


@pytest.mark.parametrize("norm", ["off", "on"])
@pytest.mark.parametrize("fc", F_CASES, ids=lambda c: f"{c[0]}{'' if c[1] is None else c[1]}@{c[2] / 1e6:g}MHz")
@pytest.mark.parametrize("win", WINDOWS, ids=lambda w: f"{w[0]}{w[1] or ''}")
@pytest.mark.parametrize("sem", ["cpu", "mex"])
def test_pw_weights_match_element_scan(oracle, sem, win, fc, norm):
    assert check_pw_case(oracle, sem, win, fc, norm) <= 1e-12


@pytest.mark.parametrize("win", WINDOWS, ids=lambda w: f"{w[0]}{w[1] or ''}")
@pytest.mark.parametrize("sem", ["cpu", "mex"])
def test_pw_normalized_weights_sum_to_one(oracle, sem, win):
    """normalizations.on (on.m:60-61): the extracted weights of a voxel sum to 1
    over the elements, and equal the un-normalized ones over their sum."""
    F = oracle.fnumber("angle_lb", 60, 1.5, 10)
    on, _, _ = extract_pw(oracle, sem, win, F, "on", 3e6)
    off, _, _ = extract_pw(oracle, sem, win, F, "off", 3e6)
    s_off = off.sum(axis=0)
    valid = np.any(off != 0.0, axis=0)
    assert valid.sum() > 0.7 * valid.size
    assert np.max(np.abs(on.sum(axis=0)[valid] - 1.0)) <= 1e-12
    assert np.max(np.abs(on[:, valid] - off[:, valid] / s_off[valid])) <= 1e-12
    assert np.all(on[:, ~valid] == 0.0)


@pytest.mark.parametrize("sem", ["cpu", "mex"])
def test_pw_weights_mirror_symmetry(oracle, sem):
    """x -> -x with the element order reversed: a_k(x) = a_{N-1-k}(-x) (the array
    is symmetric about x = 0; X holds every position with its mirror)."""
    for win in WINDOWS:
        for fc in F_CASES[:4]:
            a, _, _ = extract_pw(oracle, sem, win, _fobj(oracle, fc), "on", fc[2])
            assert np.max(np.abs(a - a[::-1, :, ::-1])) <= 1e-12


def test_pw_weights_cover_the_edge_geometries(oracle):
    """The grid holds voxels with empty apertures, voxels on element centres (d = 0)
    and on element faces, and apertures clipped by both array ends."""
    n_empty = n_centre = n_clip = 0
    for z in Z:
        for x in X:
            for F in (1.0, 3.0, 7.0):
                h = z / (2 * F)
                mem = np.flatnonzero(np.abs(XK - x) <= h)
                n_empty += mem.size == 0
                n_clip += mem.size > 0 and (mem[0] == 0 or mem[-1] == N_EL - 1)
            n_centre += np.any(XK == x)
    assert n_empty >= 10 and n_centre >= 20 and n_clip >= 10


# ---- synthetic aperture ---------------------------------------------------------

def _impulse_sa(i, k):
    rf = np.zeros((N_EL, N_EL, N_T))  # [tx, rx, t]
    rf[i, k, 0] = 1.0
    return rf


def extract_sa(O, pairs, tx, rx, sem="cpu", f_target=3e6):
    """a_tx(i) a_rx(k) per (i, k) pair of das_sa; tx/rx = (window, F, normalization)."""
    N, n1, fb = one_bin(O, sem, f_target)
    unit = (("boxcar", 0.0), O.fnumber("constant", 0.0), "off")
    out = []
    for i, k in pairs:
        rf = _impulse_sa(i, k)
        img, info = O.das_sa(X, Z, rf, FS, W, P, C0, fb, 0, tx[0], rx[0], tx[1], rx[1], tx[2], rx[2], sem)
        ref, _ = O.das_sa(X, Z, rf, FS, W, P, C0, fb, 0, unit[0], unit[0], unit[1], unit[1], "off", "off", sem)
        assert info["n_lb"] == info["n_ub"] == n1
        assert np.max(np.abs(np.abs(ref) - 2.0)) <= 1e-13
        q = img / ref
        assert np.max(np.abs(q.imag)) <= 1e-12 * max(1.0, np.max(np.abs(q.real)))
        out.append(q.real)
    return np.array(out), N, n1


SA_SIDES = [(("tukey", 0.5), ("angle_lb",), "on"), (("hann", 0.0), ("constant", 1.0), "off"),
            (("boxcar", 0.0), ("constant", 3.0), "on"), (("triangular", 0.0), ("angle_lb",), "off")]


def _sa_side(O, s):
    F = O.fnumber("angle_lb", 60, 1.5, 10) if s[1][0] == "angle_lb" else O.fnumber("constant", s[1][1])
    return (s[0], F, s[2])


@pytest.mark.parametrize("sem", ["cpu", "mex"])
@pytest.mark.parametrize("side", ["rx", "tx"])
@pytest.mark.parametrize("si", range(len(SA_SIDES)))
def test_sa_weights_match_element_scan(oracle, sem, side, si):
    """One aperture under test, the other one F = 0 / boxcar / off (all ones): the
    extracted weights are that side's alone (das_sa.m:384-414, 438-480; no depth
    clamp, das_sa.m:354-358)."""
    s = SA_SIDES[si]
    test = _sa_side(oracle, s)
    flat = (("boxcar", 0.0), oracle.fnumber("constant", 0.0), "off")
    fixed = 6
    if side == "rx":
        pairs = [(fixed, k) for k in range(N_EL)]
        a, N, n1 = extract_sa(oracle, pairs, flat, test, sem)
    else:
        pairs = [(i, fixed) for i in range(N_EL)]
        a, N, n1 = extract_sa(oracle, pairs, test, flat, sem)
    Fn = f_value(oracle, test[1], FS * n1 / N)
    for iz, z in enumerate(Z):
        for ix, x in enumerate(X):
            e = expected_weights(x, z, Fn, s[0], sem, s[2])
            if e is None:
                assert np.all(a[:, iz, ix] == 0.0)
            else:
                assert np.max(np.abs(a[:, iz, ix] - e)) <= 1e-12


@pytest.mark.parametrize("sem", ["cpu", "mex"])
def test_sa_distinct_tx_rx_weights_factor(oracle, sem):
    """Different windows, F-numbers and normalizations on the two sides: the weight
    of pair (i, k) is a_tx(i) a_rx(k) with each side normalized on its own
    (das_sa.m:479-480); the bin contributes only where both apertures are non-empty."""
    tx, rx = SA_SIDES[0], SA_SIDES[1]
    rng = np.random.default_rng(3)
    pairs = [(int(i), int(k)) for i, k in rng.integers(0, N_EL, size=(10, 2))]
    a, N, n1 = extract_sa(oracle, pairs, _sa_side(oracle, tx), _sa_side(oracle, rx), sem)
    f = FS * n1 / N
    Ftx = f_value(oracle, _sa_side(oracle, tx)[1], f)
    Frx = f_value(oracle, _sa_side(oracle, rx)[1], f)
    n_valid = 0
    for iz, z in enumerate(Z):
        for ix, x in enumerate(X):
            et = expected_weights(x, z, Ftx, tx[0], sem, tx[2])
            er = expected_weights(x, z, Frx, rx[0], sem, rx[2])
            for p, (i, k) in enumerate(pairs):
                want = 0.0 if (et is None or er is None) else et[i] * er[k]
                assert abs(a[p, iz, ix] - want) <= 1e-12
            n_valid += et is not None and er is not None
    assert n_valid > 0.5 * Z.size * X.size


def test_sa_transmit_receive_reciprocity(oracle):
    """das_sa is symmetric in its two apertures: exchanging the transmit and receive
    indices of the data together with the transmit and receive settings leaves the
    image unchanged (A9: sum_n sum_tx sum_rx a_tx a_rx e^{j w (tau_tx + tau_rx)} 2 X_rx,tx)."""
    rng = np.random.default_rng(8)
    rf = rng.standard_normal((N_EL, N_EL, N_T))
    tx = _sa_side(oracle, SA_SIDES[0])
    rx = _sa_side(oracle, SA_SIDES[2])
    # a band where angle_lb and the constant F-number both leave non-empty apertures
    fb = (2.5e6, 4.5e6)
    a, _ = oracle.das_sa(X, Z, rf, FS, W, P, C0, fb, 2, tx[0], rx[0], tx[1], rx[1], tx[2], rx[2])
    b, _ = oracle.das_sa(X, Z, np.ascontiguousarray(rf.transpose(1, 0, 2)), FS, W, P, C0, fb, 2,
                         rx[0], tx[0], rx[1], tx[1], rx[2], tx[2])
    assert np.linalg.norm(a) > 0
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)
    # and the settings do matter: keeping them while transposing the data changes the image
    c, _ = oracle.das_sa(X, Z, np.ascontiguousarray(rf.transpose(1, 0, 2)), FS, W, P, C0, fb, 2,
                         tx[0], rx[0], tx[1], rx[1], tx[2], rx[2])
    assert np.linalg.norm(a - c) > 1e-3 * np.linalg.norm(a)


# ---- the pins fail on planted mistakes ----------------------------------------------

@pytest.mark.parametrize("plant", ["OR_PLANT_CEIL_FLOOR", "OR_PLANT_SWAP_W"])
def test_planted_mistakes_are_caught(oracle, tmp_path, plant):
    """An oracle built with ceil and floor exchanged in the aperture bounds, or with
    W_L and W_R exchanged, fails the element-scan pins (the default oracle passes
    them above)."""
    lib = oracle.build_variant(str(tmp_path / plant), f"-D{plant}")
    failed = total = 0
    with oracle.using(lib):
        for sem in ("cpu", "mex"):
            for win in WINDOWS:
                for fc in F_CASES:
                    total += 1
                    try:
                        err = check_pw_case(oracle, sem, win, fc, "on")
                    except AssertionError:
                        err = np.inf
                    failed += not (err <= 1e-12)
    print(f"{plant}: {failed} of {total} cases fail")
    assert failed >= total // 4
