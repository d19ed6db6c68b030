"""ORACLE — TEST INFRASTRUCTURE ONLY: rigid-sphere Mie series.

Restates the reference's analytic oracle for the exterior rigid (Neumann)
sphere, crates/sonium-wave/src/analytical/solutions_3d.rs:315-387
(sphere_scattering_3d) and :769-820 (compute_rigid_sphere_coefficients):

    p(r, theta) = sum_n (2n+1) i^n [j_n(kr) - a_n h_n(kr)] P_n(cos theta),
    a_n = j_n'(ka) / h_n'(ka),  h_n = j_n + i y_n,

with the truncation rule suggested_sphere_terms (solutions_3d.rs:241-250),
and the sound-soft sphere (a_n = j_n(ka)/h_n(ka)) for field checks on
pressure panels.
The spherical Bessel functions are evaluated with mpmath at 40 digits rather
than by the reference's recurrences, so the coefficients are exact to double
precision; they are pinned against the reference's own 45-digit fixture
(crates/sonium-wave/fixtures/scattering-golden-v1.json) in the tests.

On the surface r = a the Wronskian j_n y_n' - j_n' y_n = 1/x^2 reduces the
bracket to i / ((ka)^2 h_n'(ka)); ``surface_pressure`` uses that form and
``total_field`` the direct series, and the tests require the two to agree.
Only tests, smoke() and bench.py's reporting may import this module.
"""

from __future__ import annotations

import math

import mpmath
import numpy as np

_DPS = 40


def suggested_sphere_terms(ka: float) -> int:
    """solutions_3d.rs:241-250: max(4*ceil(ka)+16, 17)."""
    if math.isfinite(ka) and ka > 0.0:
        return max(int(math.ceil(ka)) * 4 + 16, 17)
    return 16


def _sph(n: int, x):
    j = mpmath.sqrt(mpmath.pi / (2 * x)) * mpmath.besselj(n + mpmath.mpf(1) / 2, x)
    y = mpmath.sqrt(mpmath.pi / (2 * x)) * mpmath.bessely(n + mpmath.mpf(1) / 2, x)
    return j, y


def _sph_and_derivative(nmax: int, x):
    """j_n, y_n, j_n', y_n' for n < nmax at mpmath precision."""
    out = []
    with mpmath.workdps(_DPS):
        x = mpmath.mpf(x)
        jm1, ym1 = mpmath.cos(x) / x, mpmath.sin(x) / x  # j_{-1}, y_{-1} as in solutions_3d.rs:786-796
        for n in range(nmax):
            j, y = _sph(n, x)
            jp = jm1 - (n + 1) / x * j
            yp = ym1 - (n + 1) / x * y
            out.append((j, y, jp, yp))
            jm1, ym1 = j, y
    return out


def rigid_sphere_coefficients(ka: float, num_terms: int) -> np.ndarray:
    """a_n = j_n'(ka)/h_n'(ka) for n < num_terms (solutions_3d.rs:769-820)."""
    res = []
    with mpmath.workdps(_DPS):
        for j, y, jp, yp in _sph_and_derivative(num_terms, ka):
            a = mpmath.mpc(jp, 0) / mpmath.mpc(jp, yp)
            res.append(complex(a))
    return np.array(res, dtype=np.complex128)


def _legendre_table(nmax: int, x: np.ndarray) -> np.ndarray:
    p = np.empty((nmax,) + x.shape, dtype=np.float64)
    p[0] = 1.0
    if nmax > 1:
        p[1] = x
    for n in range(1, nmax - 1):
        p[n + 1] = ((2 * n + 1) * x * p[n] - n * p[n - 1]) / (n + 1)
    return p


def surface_coefficients(ka: float, num_terms: int) -> np.ndarray:
    """c_n = (2n+1) i^n * i / ((ka)^2 h_n'(ka)) so that p(a, theta) = sum c_n P_n."""
    res = []
    with mpmath.workdps(_DPS):
        x = mpmath.mpf(ka)
        for n, (j, y, jp, yp) in enumerate(_sph_and_derivative(num_terms, ka)):
            c = (2 * n + 1) * (mpmath.mpc(0, 1) ** n) * mpmath.mpc(0, 1) / (x * x * mpmath.mpc(jp, yp))
            res.append(complex(c))
    return np.array(res, dtype=np.complex128)


def surface_pressure(k: float, radius: float, points: np.ndarray, direction=(0.0, 0.0, 1.0),
                     num_terms: int | None = None) -> np.ndarray:
    """Total pressure on the rigid sphere surface at ``points`` (projected to r=a).

    cos(theta) is taken between the point's radial direction and the unit
    incident direction, so any plane-wave direction is supported.
    """
    ka = k * radius
    nt = num_terms or suggested_sphere_terms(ka)
    pts = np.asarray(points, dtype=np.float64)
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    ct = (pts @ d) / np.linalg.norm(pts, axis=1)
    ct = np.clip(ct, -1.0, 1.0)
    c = surface_coefficients(ka, nt)
    P = _legendre_table(nt, ct)
    return np.tensordot(c, P, axes=(0, 0))


def soft_sphere_coefficients(ka: float, num_terms: int) -> np.ndarray:
    """a_n = j_n(ka) / h_n(ka) for the sound-soft (p = 0) sphere.

    Textbook Dirichlet counterpart of the rigid coefficients (the reference's
    analytical module has no soft-sphere series; parity rests on the vanishing
    total field at r = a and on scipy's independent Bessel routines)."""
    res = []
    with mpmath.workdps(_DPS):
        for j, y, _, _ in _sph_and_derivative(num_terms, ka):
            res.append(complex(mpmath.mpc(j, 0) / mpmath.mpc(j, y)))
    return np.array(res, dtype=np.complex128)


def robin_sphere_coefficients(k: float, radius: float, sigma: complex, num_terms: int) -> np.ndarray:
    """a_n = (k j_n'(ka) - sigma j_n(ka)) / (k h_n'(ka) - sigma h_n(ka)): the
    sphere whose total field satisfies dp/dr = sigma p on r = a (a locally
    reacting surface; sigma = i rho omega Y for the admittance Y with the
    outward normal).  sigma = 0 gives the rigid coefficients."""
    res = []
    with mpmath.workdps(_DPS):
        kk = mpmath.mpf(k)
        sg = mpmath.mpc(complex(sigma).real, complex(sigma).imag)
        for j, y, jp, yp in _sph_and_derivative(num_terms, k * radius):
            res.append(complex((kk * jp - sg * j) / (kk * mpmath.mpc(jp, yp) - sg * mpmath.mpc(j, y))))
    return np.array(res, dtype=np.complex128)


def total_field(k: float, radius: float, r: float, theta: np.ndarray, num_terms: int | None = None,
                soft: bool = False, sigma: complex | None = None) -> np.ndarray:
    """Direct series sum_n (2n+1) i^n [j_n(kr) - a_n h_n(kr)] P_n(cos theta) (solutions_3d.rs:340-372);
    soft=True takes the sound-soft coefficients, sigma the Robin ones."""
    ka = k * radius
    nt = num_terms or suggested_sphere_terms(ka)
    if sigma is not None:
        a = robin_sphere_coefficients(k, radius, sigma, nt)
    else:
        a = soft_sphere_coefficients(ka, nt) if soft else rigid_sphere_coefficients(ka, nt)
    terms = []
    with mpmath.workdps(_DPS):
        for n in range(nt):
            j, y = _sph(n, mpmath.mpf(k * r))
            terms.append((complex(j), complex(mpmath.mpc(j, y))))
    ct = np.cos(np.asarray(theta, dtype=np.float64))
    P = _legendre_table(nt, ct)
    out = np.zeros(ct.shape, dtype=np.complex128)
    for n in range(nt):
        jn, hn = terms[n]
        out += (2 * n + 1) * (1j ** n) * (jn - a[n] * hn) * P[n]
    return out


def weighted_relative_l2(p: np.ndarray, p_ref: np.ndarray, weights: np.ndarray) -> float:
    """crates/sonium-qa/src/lib.rs:264-300: sqrt(sum w|p-p*|^2 / sum w|p*|^2)."""
    w = np.asarray(weights, dtype=np.float64)
    num = np.sum(w * np.abs(p - p_ref) ** 2)
    den = np.sum(w * np.abs(p_ref) ** 2)
    return float(math.sqrt(num / den))


def point_source_rigid_sphere(k: float, radius: float, source, points, strength: complex = 1.0,
                              num_terms: int = 40) -> np.ndarray:
    """Total field of a point source S e^{ikR}/(4 pi R) outside a rigid sphere.

    With e^{ikR}/(4 pi R) = (ik/4pi) h0(kR) and the addition theorem
    h0(k|x - s|) = sum_n (2n+1) j_n(k r<) h_n(k r>) P_n(cos g), the rigid
    sphere adds -(ik S/4pi) sum_n (2n+1) a_n h_n(k r_s) h_n(k r) P_n(cos g),
    a_n = j_n'(ka)/h_n'(ka) (the rigid coefficients above); g is the angle
    between x and s.  Points must satisfy a <= |x|, |x| != |s|.
    """
    s_vec = np.asarray(source, dtype=np.float64)
    rs = float(np.linalg.norm(s_vec))
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    a = rigid_sphere_coefficients(k * radius, num_terms)
    out = np.zeros(pts.shape[0], dtype=np.complex128)
    with mpmath.workdps(_DPS):
        hs = [mpmath.mpc(*_sph(n, mpmath.mpf(k * rs))) for n in range(num_terms)]
        for idx, x in enumerate(pts):
            r = float(np.linalg.norm(x))
            cg = float(np.clip(x @ s_vec / (r * rs), -1.0, 1.0))
            P = _legendre_table(num_terms, np.array([cg]))[:, 0]
            tot = mpmath.mpc(0)
            for n in range(num_terms):
                j, y = _sph(n, mpmath.mpf(k * r))
                hr = mpmath.mpc(j, y)
                if r < rs:
                    inc = mpmath.mpf(j) * hs[n]
                else:
                    jn_s = _sph(n, mpmath.mpf(k * rs))[0]
                    inc = mpmath.mpf(jn_s) * hr
                tot += (2 * n + 1) * (inc - mpmath.mpc(a[n].real, a[n].imag) * hs[n] * hr) * P[n]
            out[idx] = complex(mpmath.mpc(0, k) / (4 * mpmath.pi) * strength * tot)
    return out
