"""Pins of the oracle's analytic parameter derivatives of the stabilized system
(build_tbem_system_stabilized_parameter_derivative, tbem.rs:1858-2019, with the
derivative modes of regular.rs:92-321 and singular.rs:167-578) by central
finite differences of the oracle's own value path.

The subdivision is geometric and the self-panel rule is held at one tier, so
the finite difference sees exactly the quantity the reference differentiates;
a dropped term, a wrong sign or a missing chain-rule factor (beta, the
pressure factor, the admittance) fails these checks by orders of magnitude.
"""

import numpy as np
import pytest

import oracle
from sonium_amd.mesh import ellipsoid, icosphere

RHO, C0 = 1.21, 343.0


def _rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


def _channels(res):
    return np.array([res["g"], res["h"], res["ht"], res["e"]])


def _pair(m, i, j):
    c, nn, area = oracle.element_geometry(m.nodes, m.triangles)
    coords = m.nodes[m.triangles[j]].reshape(9)
    return c[i], nn[i], coords, area[j]


@pytest.mark.parametrize("j_offset", [1, 7, 40])
def test_regular_integral_derivative_matches_finite_difference(j_offset):
    m = icosphere(1.0, 2)
    k, h = 3.1, 3.1e-4
    src, srcn, coords, area = _pair(m, 0, j_offset)
    d = _channels(oracle.regular_integration(src, srcn, coords, area, k, derivative=True))
    fd = (_channels(oracle.regular_integration(src, srcn, coords, area, k + h))
          - _channels(oracle.regular_integration(src, srcn, coords, area, k - h))) / (2 * h)
    for ch in range(4):
        assert abs(d[ch] - fd[ch]) <= 1e-7 * max(abs(fd[ch]), 1e-3 * np.max(np.abs(fd))), ch


@pytest.mark.parametrize("params", [(3, 4, 4, 2), (6, 7, 10, 4)])
def test_singular_integral_derivative_matches_finite_difference(params):
    m = ellipsoid([1.3, 0.9, 1.0], 1)
    k, h = 2.3, 2.3e-4
    src, srcn, coords, _ = _pair(m, 5, 5)
    d = _channels(oracle.singular_integration(src, srcn, coords, k, params=params, derivative=True))
    fd = (_channels(oracle.singular_integration(src, srcn, coords, k + h, params=params))
          - _channels(oracle.singular_integration(src, srcn, coords, k - h, params=params))) / (2 * h)
    # the flat self panel has H = Ht = 0 exactly (and so do their derivatives)
    for ch in (0, 3):
        assert abs(d[ch] - fd[ch]) <= 1e-7 * abs(fd[ch]), ch
    assert abs(d[1]) < 1e-14 and abs(d[2]) < 1e-14


def _mixed_bcs(n, rng):
    bt = (rng.random(n) < 0.3).astype(np.int32)
    bv = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * (rng.random(n) < 0.5)
    adm = np.where((rng.random(n) < 0.4) & (bt == 0), 0.003 - 0.001j, np.nan)
    return bt, bv, adm


def _system(m, k, *, beta_scale, density=RHO, c=C0, **bc):
    beta = 1j / k * beta_scale
    return oracle.build_tbem_system(m.nodes, m.triangles, k, beta, density=density, speed_of_sound=c, **bc)


def test_wavenumber_derivative_matches_finite_difference():
    """build_tbem_system_stabilized_wavenumber_derivative (tbem.rs:1734-1752):
    dk/dk = 1, beta' = -i/k^2 * beta_scale, (rho omega)' = rho c."""
    m = ellipsoid([1.2, 0.8, 1.0], 1)
    n = m.num_elements
    bt, bv, adm = _mixed_bcs(n, np.random.default_rng(2))
    bc = dict(bc_type=bt, bc_value=bv, admittance=adm)
    k, h, bs = 1.9, 1.9e-4, 0.7
    A_p, b_p = _system(m, k + h, beta_scale=bs, **bc)
    A_m, b_m = _system(m, k - h, beta_scale=bs, **bc)
    dA, db = oracle.build_tbem_derivative(m.nodes, m.triangles, k, 1j / k * bs, -1j / k ** 2 * bs, **bc)
    assert _rel(dA, (A_p - A_m) / (2 * h)) < 1e-7
    assert _rel(db, (b_p - b_m) / (2 * h)) < 1e-7
    # sign flips everything
    dA2, db2 = oracle.build_tbem_derivative(m.nodes, m.triangles, k, 1j / k * bs, -1j / k ** 2 * bs, sign=-1.0,
                                            **bc)
    assert np.array_equal(dA2, -dA) and np.array_equal(db2, -db)


def test_speed_and_density_derivatives_match_finite_difference():
    """Sound speed at fixed omega (tbem.rs:1754-1782: dk/dc = -k/c, pressure
    factor fixed) and density (tbem.rs:1784-1805: only the pressure factor)."""
    m = icosphere(1.0, 1)
    n = m.num_elements
    bt, bv, adm = _mixed_bcs(n, np.random.default_rng(4))
    bc = dict(bc_type=bt, bc_value=bv, admittance=adm)
    omega, bs = 1.6 * C0, 1.0
    k = omega / C0
    hc = 1e-4 * C0
    A_p, b_p = _system(m, omega / (C0 + hc), beta_scale=bs, c=C0 + hc, **bc)
    A_m, b_m = _system(m, omega / (C0 - hc), beta_scale=bs, c=C0 - hc, **bc)
    ws = -k / C0
    dA, db = oracle.build_tbem_derivative(m.nodes, m.triangles, k, 1j / k * bs, -1j / k ** 2 * bs * ws,
                                          wavenumber_scale=ws, pressure_factor_derivative=0.0, **bc)
    assert _rel(dA, (A_p - A_m) / (2 * hc)) < 1e-7
    assert _rel(db, (b_p - b_m) / (2 * hc)) < 1e-7
    hr = 1e-4 * RHO
    A_p, b_p = _system(m, k, beta_scale=bs, density=RHO + hr, **bc)
    A_m, b_m = _system(m, k, beta_scale=bs, density=RHO - hr, **bc)
    dA, db = oracle.build_tbem_derivative(m.nodes, m.triangles, k, 1j / k * bs, 0.0, wavenumber_scale=0.0,
                                          pressure_factor_derivative=omega, **bc)
    assert _rel(dA, (A_p - A_m) / (2 * hr)) < 1e-7
    assert _rel(db, (b_p - b_m) / (2 * hr)) < 1e-7


def test_impedance_density_derivative_matches_finite_difference():
    """Y = 1/(rho c) on admittance panels (tbem.rs:1807-1830): admittance_scale = -1/rho."""
    m = icosphere(1.0, 1)
    n = m.num_elements
    k, bs = 1.3, 1.0
    omega = k * C0
    has = np.random.default_rng(6).random(n) < 0.5
    hr = 1e-4 * RHO

    def impedance(rho):
        return np.where(has, 1.0 / (rho * C0) + 0j, np.nan)

    bv = np.where(has, 0.0, 0.2 + 0.1j)
    A_p, b_p = _system(m, k, beta_scale=bs, density=RHO + hr, admittance=impedance(RHO + hr), bc_value=bv)
    A_m, b_m = _system(m, k, beta_scale=bs, density=RHO - hr, admittance=impedance(RHO - hr), bc_value=bv)
    args = (m.nodes, m.triangles, k, 1j / k * bs, 0.0)
    kw = dict(wavenumber_scale=0.0, pressure_factor_derivative=omega, admittance=impedance(RHO), bc_value=bv)
    dA, db = oracle.build_tbem_derivative(*args, admittance_scale=-1.0 / RHO, **kw)
    # Y i rho omega = i omega / c does not depend on rho: the admittance columns'
    # derivative vanishes, and only the velocity panels' right-hand side moves.
    scale = np.max(np.abs(A_p))
    assert np.max(np.abs(dA - (A_p - A_m) / (2 * hr))) < 1e-9 * scale
    assert np.max(np.abs(dA)) < 1e-12 * scale
    assert _rel(db, (b_p - b_m) / (2 * hr)) < 1e-7
    # without dY/drho the admittance columns would move by Y * d(i rho omega)
    dA0, _ = oracle.build_tbem_derivative(*args, admittance_scale=0.0, **kw)
    assert np.max(np.abs(dA0)) > 1e-4 * scale
