"""Pins of the F-number restatements (oracle.c and the product's C++ port)
against geometry the formulas must satisfy, not against retyped formulas.

angle_lb(chi, F_ub, delta) (+f_numbers/+grating/angle_lb.m:44-75): between the
branch edges the first grating lobe sits exactly at chi, i.e.
1/r - sin(theta_m) = sin(chi) with tan(theta_m) = 1/(2F) (main-lobe edge).
distance.constant(delta) (+f_numbers/+distance/constant.m:90-121): the grating
lobe stays delta away from the main-lobe edge, theta_g - theta_m = delta; in
the transition branch the grating lobe sits at 90 deg.
"""
import math

import numpy as np
import pytest

import f_number_amd as fa


def _theta_m(F):
    return math.atan(1 / (2 * F)) if F > 0 else math.pi / 2


@pytest.mark.parametrize("chi", [45.0, 60.0, 70.0])
def test_angle_lb_grating_lobe_at_chi(oracle, chi):
    F = oracle.fnumber("angle_lb", chi, 1.5, -1)  # delta < 0: no anti-aliasing term
    lo = 1 / (1 + math.sin(math.radians(chi)))
    up = 1 / (1 / math.sqrt(1 + 9) + math.sin(math.radians(chi)))
    r = np.linspace(lo, up, 41)[1:-1]
    v = oracle.fnumber_values(F, r)
    for ri, vi in zip(r, v):
        assert 1 / ri - math.sin(_theta_m(vi)) == pytest.approx(math.sin(math.radians(chi)), abs=1e-12)
    # below the lower edge: full aperture; beyond the upper edge: F_ub
    assert oracle.fnumber_values(F, np.array([0.5 * lo, lo]))[0] == 0.0
    assert list(oracle.fnumber_values(F, np.array([up, 1.5 * up]))) == [1.5, 1.5]


def test_angle_lb_reference_values(oracle):
    """angle_lb(60, 1.5): edges 0.5359 / 0.8458; 3 MHz -> 0.3809; 6.75 MHz -> 1.6287
    (delta = 10 deg term exceeds F_ub); 2.25 MHz -> 0 (full aperture)."""
    g = fa.f_numbers.grating.angle_lb(60, 1.5)
    assert g.lower_bound == pytest.approx(0.5359, abs=1e-4)
    assert g.upper_bound == pytest.approx(0.8458, abs=1e-4)
    F = oracle.fnumber("angle_lb", 60, 1.5, 10)
    r = 0.3048e-3 * np.array([3e6, 6.75e6, 2.25e6]) / 1540
    v = oracle.fnumber_values(F, r)
    assert v[0] == pytest.approx(0.3809, abs=1e-4)
    assert v[1] == pytest.approx(1.6287, abs=1e-4)
    assert v[2] == 0.0


@pytest.mark.parametrize("delta", [5.0, 10.0, 20.0])
def test_distance_constant_keeps_angular_distance(oracle, delta):
    F = oracle.fnumber("distance_constant", delta, np.inf)
    o = 1 + math.cos(math.radians(delta))
    s = math.sin(math.radians(delta))
    r = np.linspace(1 / o, 1 / s, 60)[1:-1]
    v = oracle.fnumber_values(F, r)
    for ri, vi in zip(r, v):
        th_m = _theta_m(vi)
        assert math.degrees(math.asin(1 / ri - math.sin(th_m)) - th_m) == pytest.approx(delta, abs=1e-9)
    # transition branch 0.5 <= r < 1/o: grating lobe at 90 deg
    rt = np.linspace(0.5, 1 / o, 7)[1:-1]
    for ri, vi in zip(rt, oracle.fnumber_values(F, rt)):
        assert 1 / ri - math.sin(_theta_m(vi)) == pytest.approx(1.0, abs=1e-12)
    assert oracle.fnumber_values(F, np.array([0.3]))[0] == 0.0
    assert oracle.fnumber_values(oracle.fnumber("distance_constant", delta, 2.0), np.array([1.5 / s]))[0] == 2.0


@pytest.mark.parametrize("kind,soft,att", [("directivity_soft", True, 3.0), ("directivity_rigid", False, 6.0),
                                           ("perrot", True, 3.0)])
def test_directivity_threshold_crossing(oracle, kind, soft, att):
    """F = 1/(2 tan alpha) where the element pattern O(theta) sinc(w sin theta) equals
    10^(-att/20); fminbnd's TolX = 1e-4 bounds alpha (the iterates are parity unpinned)."""
    w_over_p = 0.918
    F = oracle.fnumber(kind, w_over_p, att)
    r = np.array([0.6, 0.9, 1.3])
    v = oracle.fnumber_values(F, r)
    for ri, vi in zip(r, v):
        alpha = math.atan(1 / (2 * vi))
        w = w_over_p * ri

        def pat(th):
            x = w * math.sin(th)
            sinc = 1.0 if x == 0 else math.sin(math.pi * x) / (math.pi * x)
            return (math.cos(th) if soft else 1.0) * sinc

        # exact crossing by bisection
        lo, hi = 0.0, math.pi / 2
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if pat(mid) > 10 ** (-att / 20):
                lo = mid
            else:
                hi = mid
        assert abs(alpha - lo) < 2e-4


def test_szabo_constant_none(oracle):
    r = np.array([0.4, 0.8, 1.6])
    assert np.allclose(oracle.fnumber_values(oracle.fnumber("szabo", 0.918), r), 0.918 * r / 1.2, rtol=0, atol=1e-15)
    assert np.all(oracle.fnumber_values(oracle.fnumber("constant", 1.5), r) == 1.5)
    # distance.none = distance.constant(0, inf): grating lobe on the main-lobe edge
    r2 = np.array([0.4, 0.6, 2.0])
    none = oracle.fnumber_values(oracle.fnumber("distance_none"), r2)
    assert none[0] == 0.0
    for ri, vi in zip(r2[1:], none[1:]):
        assert 1 / ri - math.sin(_theta_m(vi)) == pytest.approx(math.sin(_theta_m(vi)), abs=1e-12)


KINDS = [
    (fa.f_numbers.grating.angle_lb(60, 1.5), ("angle_lb", 60, 1.5, 10)),
    (fa.f_numbers.grating.angle_lb(70, 1.5, 10), ("angle_lb", 70, 1.5, 10)),
    (fa.f_numbers.grating.angle_lb(45, 3.0, -1), ("angle_lb", 45, 3.0, -1)),
    (fa.f_numbers.distance.constant(10), ("distance_constant", 10, np.inf)),
    (fa.f_numbers.distance.constant(5, 2.5), ("distance_constant", 5, 2.5)),
    (fa.f_numbers.distance.none(), ("distance_none",)),
    (fa.f_numbers.constant(1.5), ("constant", 1.5)),
    (fa.f_numbers.directivity.szabo(0.918), ("szabo", 0.918)),
    (fa.f_numbers.directivity.soft(0.918, 3), ("directivity_soft", 0.918, 3)),
    (fa.f_numbers.directivity.rigid(0.918, 6), ("directivity_rigid", 0.918, 6)),
    (fa.f_numbers.directivity.perrot(0.918), ("perrot", 0.918)),
]


@pytest.mark.parametrize("prod,orc", KINDS, ids=[str(k[0]) for k in KINDS])
def test_product_port_matches_oracle(oracle, prod, orc):
    """The product's C++ port (fn_f_number_values) equals the oracle bit for bit."""
    r = np.linspace(0.05, 2.0, 157)
    got = prod.compute_values(r)
    ref = oracle.fnumber_values(oracle.fnumber(*orc), r)
    assert np.array_equal(got, ref)


def test_f_number_derived_quantities():
    """f_number.m:119-298: main-lobe bound, grating distance, depth of field."""
    F = fa.f_numbers.constant(1.0)
    r = np.array([0.5, 1.0])
    assert np.allclose(F.compute_bounds_main(r), 1 / math.sqrt(5))
    assert np.allclose(F.compute_distances_grating(r), 1 / r - 1 / math.sqrt(5))
    assert np.allclose(F.compute_depths_of_field(r), 6.1434 / r)


def test_f_number_rejects_nonpositive_ratio():
    with pytest.raises(fa.FNumberError) as e:
        fa.f_numbers.constant(1.0).compute_values(np.array([0.5, 0.0]))
    assert e.value.identifier.startswith("compute_values:")
