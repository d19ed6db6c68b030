"""Pins of the oracle's quadrature tables (gauss.rs:14-154, 199-540).

Independent checks: numpy's Gauss–Legendre nodes, exact polynomial moments
on the reference triangle, and the reference's own unit tests
(gauss.rs:546-653).
"""

import math

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("order", list(range(1, 13)))
def test_gauss_legendre_matches_numpy(order):
    x, w = oracle.gauss_legendre(order)
    xr, wr = np.polynomial.legendre.leggauss(order)
    # tables carry 16 significant digits; generated orders are Newton-exact
    assert np.max(np.abs(x - xr)) < 2e-15
    assert np.max(np.abs(w - wr)) < 2e-15


@pytest.mark.parametrize("order", [3, 5, 7])
def test_generated_orders_are_not_rounded(order):
    """gauss.rs:590-611: generated rules keep their order, sorted, weights sum to 2."""
    x, w = oracle.gauss_legendre(order)
    assert len(x) == order
    assert np.all(np.diff(x) > 0)
    assert abs(w.sum() - 2.0) < 2e-14
    # exact for degree 2n-1, not for 2n
    deg = 2 * order - 2
    assert abs(np.sum(w * x ** deg) - 2.0 / (deg + 1)) < 1e-14
    deg2 = 2 * order
    assert abs(np.sum(w * x ** deg2) - 2.0 / (deg2 + 1)) > 1e-6


def _monomial_exact(a, b):
    """Integral of x^a y^b over the unit triangle = a! b! / (a+b+2)!."""
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("order,degree,npts", [(1, 1, 1), (2, 3, 4), (3, 5, 7), (4, 7, 13)])
def test_triangle_rules_are_exact_to_their_degree(order, degree, npts):
    q = oracle.triangle_quadrature(order)
    assert q.shape[0] == npts
    # the tables carry 15 significant digits
    assert abs(q[:, 2].sum() - 0.5) < 5e-15
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            got = np.sum(q[:, 2] * q[:, 0] ** a * q[:, 1] ** b)
            assert abs(got - _monomial_exact(a, b)) < 2e-15, (a, b)


def test_tr13_is_not_exact_beyond_degree_7():
    q = oracle.triangle_quadrature(4)
    errs = [abs(np.sum(q[:, 2] * q[:, 0] ** a * q[:, 1] ** (8 - a)) - _monomial_exact(a, 8 - a)) for a in range(9)]
    assert max(errs) > 1e-8


def test_collapsed_rule_above_order_four():
    """gauss.rs:628-643: order 12 is a 144-point Duffy-collapsed rule, x^10 y^8 exact."""
    q = oracle.triangle_quadrature(12)
    assert q.shape[0] == 144
    assert abs(q[:, 2].sum() - 0.5) < 1e-14
    got = np.sum(q[:, 2] * q[:, 0] ** 10 * q[:, 1] ** 8)
    assert abs(got - _monomial_exact(10, 8)) < 1e-15
