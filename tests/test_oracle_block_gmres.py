"""Pins of the oracle's block GMRES (block_gmres.rs:163-215, 463-1120).

The reference's own tests of the block method (block_gmres.rs:1268-1993) on
its fixtures — the 1-D Laplacian, its complex-shifted variant, duplicated,
scaled and zero right-hand-side columns, restart 1 — are restated here with
their assertions; then LAPACK on random complex systems, the Jacobi-
preconditioned path, and two properties of the method: with one right-hand
side the block recurrence is scalar GMRES (same step count and solution),
and one shared block Krylov space K_m(A, R0) contains every column's own
K_m(A, r0_j), so without restarts no column needs more block steps than its
own scalar solve.
"""

import numpy as np
import pytest

import oracle


def laplacian_1d(n, diag=2.0):
    A = np.zeros((n, n), dtype=np.complex128)
    for i in range(n):
        A[i, i] = diag
        if i > 0:
            A[i, i - 1] = -1.0
        if i < n - 1:
            A[i, i + 1] = -1.0
    return A


def rel_res(A, X, B, j):
    return np.linalg.norm(B[:, j] - A @ X[:, j]) / np.linalg.norm(B[:, j])


def test_single_rhs():  # block_gmres.rs:1268-1288
    n = 50
    A = laplacian_1d(n)
    B = (np.arange(n) + 1.0).reshape(n, 1).astype(np.complex128)
    X, st = oracle.block_gmres(A, B, restart=30, max_iterations=50, tol=1e-10)
    assert st["converged"]


def test_restart_one_performs_an_arnoldi_step():  # :1290-1319
    A = np.eye(2, dtype=np.complex128)
    B = np.array([[2.0], [-3.0]], dtype=np.complex128)
    X, st = oracle.block_gmres(A, B, restart=1, max_iterations=1, tol=1e-12)
    assert st["converged"] and st["iterations"] == 1 and st["true_residuals"][0] < 1e-12
    assert abs(X[0, 0] - 2.0) < 1e-12 and abs(X[1, 0] + 3.0) < 1e-12


def _three_rhs(n):
    i = np.arange(n)
    return np.stack([np.ones(n), np.sin(i), np.where(i < n // 2, 1.0, -1.0)], 1).astype(np.complex128)


def test_multiple_rhs_and_true_residuals():  # :1347-1381, 1817-1863
    n = 40
    A = laplacian_1d(n)
    B = _three_rhs(n)
    X, st = oracle.block_gmres(A, B, restart=40, max_iterations=50, tol=1e-10)
    assert st["converged"]
    for j in range(3):
        direct = rel_res(A, X, B, j)
        assert abs(st["true_residuals"][j] - direct) < 1e-14
        assert st["true_residuals"][j] <= st["residuals"][j] * 10.0 + 1e-12
        assert np.isfinite(st["backward_errors"][j])


def test_rank_deficient_block():  # :1516-1556
    n = 40
    A = laplacian_1d(n)
    i = np.arange(n)
    c0 = np.sqrt(i + 1.0)
    B = np.stack([c0, c0, np.cos(i * 0.2)], 1).astype(np.complex128)
    X, st = oracle.block_gmres(A, B, restart=30, max_iterations=50, tol=1e-10)
    assert st["converged"]
    assert np.max(np.abs(X[:, 0] - X[:, 1])) < 1e-8
    assert all(rel_res(A, X, B, j) < 1e-9 for j in range(3))


def test_zero_rhs_column():  # :1558-1597
    n = 32
    A = laplacian_1d(n)
    i = np.arange(n)
    B = np.stack([np.zeros(n), np.sin(i * 0.31) + 0.25], 1).astype(np.complex128)
    X, st = oracle.block_gmres(A, B, restart=24, max_iterations=50, tol=1e-10)
    assert st["converged"]
    assert np.all(X[:, 0] == 0.0) and st["true_residuals"][0] < 1e-12
    assert rel_res(A, X, B, 1) < 1e-9


def test_scaled_dependent_rhs_columns():  # :1599-1635
    n = 36
    A = laplacian_1d(n)
    i = np.arange(n)
    base = np.cos(i * 0.23) + 0.4
    B = np.stack([base, 2.0 * base, np.sin(i * 0.47) - 0.1], 1).astype(np.complex128)
    X, st = oracle.block_gmres(A, B, restart=24, max_iterations=50, tol=1e-10)
    assert st["converged"]
    assert np.max(np.abs(X[:, 1] - 2.0 * X[:, 0])) < 1e-8
    assert all(rel_res(A, X, B, j) < 1e-9 for j in range(3))


def test_complex_shifted_laplacian():  # :1637-1669
    n = 40
    A = laplacian_1d(n, diag=2.0 + 0.5j)
    i = np.arange(n)
    B = np.stack([np.sin(i) + 1j, 1.0 + 1j * np.cos(i * 0.1)], 1)
    X, st = oracle.block_gmres(A, B, restart=40, max_iterations=50, tol=1e-8)
    assert st["converged"] and all(rel_res(A, X, B, j) < 1e-7 for j in range(2))


def test_guess_reduces_iterations():  # :1755-1815
    n = 50
    A = laplacian_1d(n, diag=6.0)
    i = np.arange(n)
    B = np.stack([np.ones(n), np.sin(i)], 1).astype(np.complex128)
    X, cold = oracle.block_gmres(A, B, restart=40, max_iterations=50, tol=1e-8)
    assert cold["converged"] and cold["iterations"] > 0
    _, exact = oracle.block_gmres(A, B, X0=X, restart=40, max_iterations=50, tol=1e-8)
    assert exact["converged"] and exact["iterations"] == 0
    Xw, warm = oracle.block_gmres(A, B, X0=X * 1.001, restart=40, max_iterations=50, tol=1e-8)
    assert warm["converged"] and warm["iterations"] < cold["iterations"]


def test_converged_gates_on_configured_residual():  # :1865-1911
    n = 40
    A = laplacian_1d(n)
    i = np.arange(n)
    B = np.stack([np.ones(n), np.cos(i * 0.2)], 1).astype(np.complex128)
    for trc in (True, False):
        X, st = oracle.block_gmres(A, B, restart=40, max_iterations=50, tol=1e-10, true_residual_check=trc)
        gated = st["true_residuals"] if trc else st["residuals"]
        assert st["converged"] == bool(np.all(gated <= 1e-10)) and st["converged"]
    _, st = oracle.block_gmres(A, B, restart=40, max_iterations=0, tol=1e-10, true_residual_check=False)
    assert not st["converged"] and np.all(st["true_residuals"] > 1e-10)


def test_restart_one_across_cycles():  # :1950-1992
    n = 25
    A = np.diag(1.0 + (np.arange(n) % 5) * 0.5).astype(np.complex128)
    B = np.stack([np.ones(n), 2.0 * np.ones(n)], 1).astype(np.complex128)
    X, st = oracle.block_gmres(A, B, restart=1, max_iterations=100, tol=1e-8)
    assert st["iterations"] > 1 and st["converged"]
    assert all(rel_res(A, X, B, j) < 1e-6 for j in range(2))


def _random_system(n, s, seed):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 3.0 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    B = rng.standard_normal((n, s)) + 1j * rng.standard_normal((n, s))
    return A.astype(np.complex128), B


@pytest.mark.parametrize("jacobi", [False, True])
def test_matches_lapack(jacobi):
    A, B = _random_system(120, 8, 4)
    if jacobi:
        A = A * (1.0 + np.arange(120) / 10.0)[:, None]  # row scaling the inverse diagonal undoes
    X, st = oracle.block_gmres(A, B, restart=30, max_iterations=20, tol=1e-10, jacobi=jacobi)
    assert st["converged"]
    Xd = np.linalg.solve(A, B)
    assert np.linalg.norm(X - Xd) / np.linalg.norm(Xd) < 1e-8


def test_single_column_is_scalar_gmres():
    A, B = _random_system(150, 1, 7)
    X, st = oracle.block_gmres(A, B, restart=60, max_iterations=5, tol=1e-10)
    x, sc = oracle.gmres(A, B[:, 0], restart=60, max_iterations=5, tol=1e-10)
    assert st["converged"] and sc["converged"]
    assert abs(st["iterations"] - sc["iterations"]) <= 1
    assert np.linalg.norm(X[:, 0] - x) / np.linalg.norm(x) < 1e-9


def test_shared_basis_needs_no_more_steps_than_any_column():
    n, s = 200, 6
    rng = np.random.default_rng(9)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    d = np.exp(1j * rng.uniform(0, 0.8, n)) * np.linspace(1.0, 40.0, n)
    A = (Q * d) @ Q.conj().T
    B = rng.standard_normal((n, s)) + 1j * rng.standard_normal((n, s))
    X, st = oracle.block_gmres(A, B, restart=150, max_iterations=1, tol=1e-9)
    steps = [oracle.gmres(A, B[:, j], restart=150, max_iterations=1, tol=1e-9)[1]["iterations"] for j in range(s)]
    assert st["converged"]
    assert st["iterations"] <= max(steps), (st["iterations"], steps)
    assert st["iterations"] < min(steps), "the shared space should beat every single column here"
