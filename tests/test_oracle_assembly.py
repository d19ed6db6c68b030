"""Pins of the oracle's assembled system and solver (tbem.rs:1284-1466,
incident.rs:807-834, gmres.rs:631-1057), independent of the restatement:

* Gauss's solid-angle identity: with beta = 0 and k -> 0 every row of
  A = -I/2 + H sums to -1 (the reference diagnoses the same row sums,
  tbem.rs:1668-1700).  On a closed polyhedron the double layer of a constant
  is exactly -1/2 at a point of a flat face, so the sum is exact up to
  quadrature error, at every refinement level;
* the hypersingular operator annihilates constants at k = 0 (rows of E sum to
  0, exactly on a closed polyhedron up to quadrature error);
* partial rows equal the full assembly and results do not depend on the
  thread count (tbem.rs:5619, :6191);
* GMRES against LAPACK (numpy.linalg.solve) and its own recomputed residual;
* end to end against the Mie series, including the reference's surface gate
  (0.5 % at ka = 1 with beta = 0.2 i/k, bem_mie_validation.rs:18, 298-386).
"""

import numpy as np
import pytest

import oracle
from oracle import mie
from sonium_amd.mesh import icosphere


@pytest.mark.parametrize("subdiv", [1, 2, 3])
def test_gauss_identity_row_sums(subdiv):
    m = icosphere(1.0, subdiv)
    A, _ = oracle.build_tbem_system(m.nodes, m.triangles, 1e-6, 0.0)
    rs = A.sum(axis=1)
    assert np.max(np.abs(rs + 1.0)) < 1e-7
    # with beta = 0 the diagonal is the free term alone: H_ii = 0 on a flat
    # panel (singular.rs:1763-1800), so a sign error there is caught too
    assert np.max(np.abs(np.diag(A) + 0.5)) < 1e-13


def test_hypersingular_annihilates_constants():
    for s in (2, 3):
        m = icosphere(1.0, s)
        A0, _ = oracle.build_tbem_system(m.nodes, m.triangles, 1e-6, 0.0)
        A1, _ = oracle.build_tbem_system(m.nodes, m.triangles, 1e-6, 1.0)
        E = A1 - A0  # beta = 1 isolates E
        scale = np.max(np.abs(np.diag(E)))
        assert np.max(np.abs(E.sum(axis=1))) / scale < 1e-3


def test_partial_rows_and_thread_independence():
    m = icosphere(1.0, 2)
    k = 2.0
    A, b = oracle.build_tbem_system(m.nodes, m.triangles, k, 1j / k, threads=1)
    A8, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, 1j / k, threads=8)
    assert np.array_equal(A, A8)
    Ap, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, 1j / k, rows=(37, 101), threads=3)
    assert np.array_equal(Ap, A[37:101])


def test_free_term_and_beta_linearity():
    """A(beta) = A(0) + beta*E: the matrix is affine in beta (tbem.rs:1405)."""
    m = icosphere(1.0, 1)
    k = 1.5
    A0, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, 0.0)
    A1, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, 1.0)
    Ab, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, 0.3 - 0.7j)
    assert np.max(np.abs(Ab - (A0 + (0.3 - 0.7j) * (A1 - A0)))) < 1e-13 * np.max(np.abs(A1))


def _c1():
    m = icosphere(1.0, 3)
    k = 2.0
    beta = 1j / k
    A, r = oracle.build_tbem_system(m.nodes, m.triangles, k, beta)
    c, n, a = m.element_geometry()
    b = oracle.plane_wave_rhs(c, n, k, beta, rhs=r.copy())
    return m, A, b, c, a


def test_gmres_against_lapack():
    m, A, b, c, a = _c1()
    x, st = oracle.gmres(A, b, restart=50, max_iterations=100, tol=1e-8)
    xl = np.linalg.solve(A, b)
    assert st["converged"] and st["status"] == "converged"
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-8
    assert abs(st["residual"] - np.linalg.norm(b - A @ x) / np.linalg.norm(b)) < 1e-12
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) < 1e-7
    # restarted and preconditioned variants converge to the same solution
    x2, st2 = oracle.gmres(A, b, restart=7, max_iterations=200, tol=1e-10, jacobi=True)
    assert st2["converged"] and st2["restarts"] > 0
    assert np.linalg.norm(x2 - xl) / np.linalg.norm(xl) < 1e-8


def test_gmres_invalid_and_trivial_inputs():
    A = np.eye(4, dtype=complex) * 2.0
    x, st = oracle.gmres(A, np.array([1, 2, np.nan, 0], dtype=complex))
    assert st["status"] == "numerical_error"
    x, st = oracle.gmres(A, np.zeros(4, dtype=complex))
    assert st["converged"] and np.all(x == 0)
    x, st = oracle.gmres(A, np.ones(4, dtype=complex), restart=2, tol=1e-12)
    assert st["converged"] and np.allclose(x, 0.5)


def test_c1_surface_pressure_vs_mie():
    m, A, b, c, a = _c1()
    x = np.linalg.solve(A, b)
    err = mie.weighted_relative_l2(x, mie.surface_pressure(2.0, 1.0, c), a)
    assert 1e-3 < err < 1e-2  # beta = i/k on 1280 panels (discretisation error ~0.57 %)


def test_reference_mie_gate_ka1():
    """bem_mie_validation.rs:18, 298-386: < 0.5 % at ka = 1, beta = 0.2 i/k, 1280 panels."""
    m = icosphere(1.0, 3)
    k = 1.0
    beta = 0.2j / k
    A, r = oracle.build_tbem_system(m.nodes, m.triangles, k, beta)
    c, n, a = m.element_geometry()
    b = oracle.plane_wave_rhs(c, n, k, beta, rhs=r)
    x = np.linalg.solve(A, b)
    assert mie.weighted_relative_l2(x, mie.surface_pressure(k, 1.0, c), a) < 0.005


def test_mesh_refinement_reduces_mie_error():
    errs = []
    for s in (2, 3):
        m = icosphere(1.0, s)
        k = 2.0
        beta = 1j / k
        A, r = oracle.build_tbem_system(m.nodes, m.triangles, k, beta)
        c, n, a = m.element_geometry()
        b = oracle.plane_wave_rhs(c, n, k, beta, rhs=r)
        x = np.linalg.solve(A, b)
        errs.append(mie.weighted_relative_l2(x, mie.surface_pressure(k, 1.0, c), a))
    assert errs[1] < 0.5 * errs[0]


def test_plane_wave_direction_rotates_solution():
    """A rotated incident direction gives the Mie field about that direction."""
    m = icosphere(1.0, 3)
    k = 1.0
    beta = 1j / k
    A, r = oracle.build_tbem_system(m.nodes, m.triangles, k, beta)
    c, n, a = m.element_geometry()
    d = np.array([1.0, -1.0, 0.5])
    d /= np.linalg.norm(d)
    b = oracle.plane_wave_rhs(c, n, k, beta, direction=d, rhs=r)
    x = np.linalg.solve(A, b)
    assert mie.weighted_relative_l2(x, mie.surface_pressure(k, 1.0, c, direction=d), a) < 0.01
