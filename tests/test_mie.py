"""Pins of the Mie oracle (solutions_3d.rs:241-387, 769-820) against the
reference's own fixtures:

* crates/sonium-wave/fixtures/scattering-golden-v1.json — rigid-sphere a_n at
  ka = 1 (20 terms) and ka = 5 (36 terms), 45 digits (committed under
  tests/golden/);
* crates/sonium-wave/fixtures/special-golden-wolfram-v1.json — j_n, y_n at
  x in {0.1, pi/4, pi/2, pi, 2pi, 50} from Wolfram (tests/golden/);
and against scipy.special (an independent implementation) at ka = 2, 10, 14,
plus the Wronskian identity that the surface formula uses.
"""

import json
import math
import os

import mpmath
import numpy as np
import pytest
from scipy import special

from oracle import mie

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_rigid_coefficients_match_reference_golden():
    g = _load("scattering-golden-v1.json")
    for case in g["sphere"]:
        ka = float(case["ka"])
        ref = np.array([complex(float(c["re"]), float(c["im"])) for c in case["coefficients"]])
        got = mie.rigid_sphere_coefficients(ka, len(ref))
        # the fixture is rounded to 45 decimal places (absolute precision 5e-46)
        err = np.abs(got - ref)
        assert np.all(err <= 1e-13 * np.abs(ref) + 1e-45), (ka, err.max())


def test_spherical_bessel_match_wolfram_golden():
    g = _load("special-golden-wolfram-v1.json")
    for case in g["spherical"]:
        js = [float(v) for v in case["j"]]
        ys = [float(v) for v in case["y"]]
        with mpmath.workdps(60):  # the fixture's x carries 50 digits (pi/2 exactly, ...)
            x = mpmath.mpf(str(case["x"]))
            vals = [mie._sph(n, x) for n in range(len(js))]
        for n, ((j, y), jr, yr) in enumerate(zip(vals, js, ys)):
            # absolute 1e-40: y_0(pi/2) is exactly 0 in the 50-digit fixture
            assert abs(float(j) - jr) <= 1e-14 * abs(jr) + 1e-40, (case["name"], n)
            assert abs(float(y) - yr) <= 1e-14 * abs(yr) + 1e-40, (case["name"], n)


@pytest.mark.parametrize("ka", [2.0, 10.0, 14.0])
def test_coefficients_against_scipy(ka):
    n = mie.suggested_sphere_terms(ka)
    a = mie.rigid_sphere_coefficients(ka, n)
    nn = np.arange(n)
    jp = special.spherical_jn(nn, ka, derivative=True)
    yp = special.spherical_yn(nn, ka, derivative=True)
    ref = jp / (jp + 1j * yp)
    big = np.abs(ref) > 1e-12
    assert np.max(np.abs(a[big] - ref[big]) / np.abs(ref[big])) < 1e-10


@pytest.mark.parametrize("ka", [1.0, 2.0, 10.0])
def test_surface_wronskian_form_equals_direct_series(ka):
    theta = np.linspace(0.0, math.pi, 13)
    direct = mie.total_field(ka, 1.0, 1.0, theta)
    pts = np.stack([np.sin(theta), np.zeros_like(theta), np.cos(theta)], axis=1)
    wr = mie.surface_pressure(ka, 1.0, pts)
    assert np.max(np.abs(wr - direct)) < 1e-11 * np.max(np.abs(direct))


def test_truncation_rule_and_convergence():
    assert mie.suggested_sphere_terms(1.0) == 20
    assert mie.suggested_sphere_terms(10.0) == 56
    assert mie.suggested_sphere_terms(0.1) == 20
    assert mie.suggested_sphere_terms(float("nan")) == 16
    pts = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    p1 = mie.surface_pressure(10.0, 1.0, pts)
    p2 = mie.surface_pressure(10.0, 1.0, pts, num_terms=80)
    assert np.max(np.abs(p1 - p2)) < 1e-12


def test_rigid_sphere_low_frequency_limit():
    """ka -> 0: the total surface pressure tends to the incident field plus the
    rigid-sphere dipole/monopole corrections; at ka = 1e-3 it is ~1 everywhere."""
    pts = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 0.0, -1.0]])
    p = mie.surface_pressure(1e-3, 1.0, pts)
    assert np.max(np.abs(p - 1.0)) < 5e-3


@pytest.mark.parametrize("ka", [0.5, 1.0, 5.0])
def test_soft_sphere_series(ka):
    """Sound-soft sphere: the total field vanishes on r = a, the coefficients
    match scipy's spherical Bessel functions, and off the sphere the series
    differs from the rigid one."""
    n = mie.suggested_sphere_terms(ka)
    a = mie.soft_sphere_coefficients(ka, n)
    nn = np.arange(n)
    j = special.spherical_jn(nn, ka)
    y = special.spherical_yn(nn, ka)
    ref = j / (j + 1j * y)
    big = np.abs(ref) > 1e-12
    assert np.max(np.abs(a[big] - ref[big]) / np.abs(ref[big])) < 1e-10
    theta = np.linspace(0.0, math.pi, 9)
    on = mie.total_field(ka, 1.0, 1.0, theta, soft=True)
    assert np.max(np.abs(on)) < 1e-12
    off_soft = mie.total_field(ka, 1.0, 2.0, theta, soft=True)
    off_rigid = mie.total_field(ka, 1.0, 2.0, theta)
    assert np.max(np.abs(off_soft - off_rigid)) > 1e-2


def test_robin_sphere_series():
    """dp/dr = sigma p on r = a (central difference of the series in r), and
    sigma = 0 reproduces the rigid coefficients."""
    k, a = 1.3, 1.0
    n = mie.suggested_sphere_terms(k * a)
    assert np.allclose(mie.robin_sphere_coefficients(k, a, 0.0, n), mie.rigid_sphere_coefficients(k * a, n),
                       rtol=1e-13, atol=1e-16)
    sigma = 0.25 + 0.5j
    theta = np.linspace(0.0, math.pi, 7)
    h = 1e-5
    plus = mie.total_field(k, a, a + h, theta, sigma=sigma)
    minus = mie.total_field(k, a, a - h, theta, sigma=sigma)
    on = mie.total_field(k, a, a, theta, sigma=sigma)
    dpdr = (plus - minus) / (2 * h)
    assert np.max(np.abs(dpdr - sigma * on)) < 1e-8 * np.max(np.abs(on))


def test_point_source_series_reduces_to_free_field_and_is_rigid():
    """Without the sphere term the series is the free-field monopole; with it,
    dp/dr = 0 on r = a (central difference)."""
    k, a = 1.1, 1.0
    s = np.array([0.0, 0.0, 3.0])
    x = np.array([[0.5, 0.2, 1.4], [-1.6, 0.3, 0.2]])
    R = np.linalg.norm(x - s, axis=1)
    free = np.exp(1j * k * R) / (4 * np.pi * R)
    big = mie.point_source_rigid_sphere(k, 1e-6, s, x)  # a vanishing sphere scatters nothing
    assert np.max(np.abs(big - free)) < 1e-10 * np.max(np.abs(free))
    u = np.array([[0.6, 0.0, 0.8], [0.0, -1.0, 0.0]])
    h = 1e-5
    pp = mie.point_source_rigid_sphere(k, a, s, u * (a + h))
    pm = mie.point_source_rigid_sphere(k, a, s, u * (a - h))
    on = mie.point_source_rigid_sphere(k, a, s, u * a)
    assert np.max(np.abs(pp - pm) / (2 * h)) < 1e-7 * np.max(np.abs(on))
