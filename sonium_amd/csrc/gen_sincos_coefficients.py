"""Coefficients of the device sin/cos polynomials on |r| <= pi/2.

sin r = r + r^3 P(r^2),   cos r = 1 - r^2/2 + r^4 Q(r^2),

with P and Q of degree 7 interpolating (sin r - r)/r^3 and
(cos r - 1 + r^2/2)/r^4 as functions of u = r^2 at the eight Chebyshev nodes
of [0, (pi/2)^2] (Chebyshev interpolation, near-minimax for these entire
functions), computed in 50-digit arithmetic with mpmath and rounded to double.
The device reduces its argument by pi, so one sign flip finishes both values.
``python gen_sincos_coefficients.py`` prints the constants and the maximum
error of the polynomials with the rounded coefficients, measured against
mpmath on 2001 points of [0, pi/2].
"""

from __future__ import annotations

import mpmath as mp

DEGREE = 7


def _half_range():
    return mp.pi / 2  # evaluated at the working precision


def _g(u):
    if u == 0:
        return mp.mpf(-1) / 6
    r = mp.sqrt(u)
    return (mp.sin(r) - r) / r ** 3


def _h(u):
    if u == 0:
        return mp.mpf(1) / 24
    r = mp.sqrt(u)
    return (mp.cos(r) - 1 + u / 2) / u ** 2


def _interpolate(f, degree):
    upper = _half_range() ** 2
    n = degree + 1
    nodes = [upper / 2 * (1 - mp.cos((2 * i + 1) * mp.pi / (2 * n))) for i in range(n)]
    vander = mp.matrix([[x ** j for j in range(n)] for x in nodes])
    values = mp.matrix([f(x) for x in nodes])
    return [float(c) for c in mp.lu_solve(vander, values)]


def coefficients():
    """(P, Q) as lists of doubles, lowest degree first."""
    with mp.workdps(50):
        return _interpolate(_g, DEGREE), _interpolate(_h, DEGREE)


def max_errors(P, Q, samples=2001):
    """Maximum |poly - exact| of sin and cos over [0, pi/2] with the given coefficients."""
    es = ec = mp.mpf(0)
    with mp.workdps(50):
        for i in range(samples):
            r = _half_range() * i / (samples - 1)
            u = r * r
            p = sum(mp.mpf(c) * u ** j for j, c in enumerate(P))
            q = sum(mp.mpf(c) * u ** j for j, c in enumerate(Q))
            es = max(es, abs(r + r ** 3 * p - mp.sin(r)))
            ec = max(ec, abs(1 - u / 2 + u * u * q - mp.cos(r)))
    return float(es), float(ec)


if __name__ == "__main__":
    P, Q = coefficients()
    print("P =", ", ".join(repr(c) for c in P))
    print("Q =", ", ".join(repr(c) for c in Q))
    print("max error sin %.3e cos %.3e" % max_errors(P, Q))
