"""Pins of the GCRO-DR oracle (oracle/gcro_dr.py), CPU only.

- the small eigensolver: on Hessenberg input (where the reference's reduction
  is an exact similarity) its eigenvalues equal LAPACK's and every returned
  pair satisfies H y = theta y;
- the reduction itself: Q is unitary and Q^H A Q equals the returned matrix on
  and above the subdiagonal; for real data nothing is lost below it, for
  complex data the zeroed part is not small (the reference's behaviour, kept);
- the reference's own GCRO-DR tests (gcro_dr.rs:1114-1400): first solve
  converges and seeds the subspace, recycling does not slow a nearby solve,
  same solution as GMRES/LAPACK, a 60-frequency sweep at least halves the
  total Arnoldi steps, the diagonal smoke case, the restart after an
  optimistic exact gate, and no subspace change when a solve fails.
"""

import numpy as np
import pytest

from oracle import gcro_dr as G


def helmholtz_like(n, freq):
    """build_helmholtz_like (gcro_dr.rs:1097-1112)."""
    k2 = freq * freq * 0.05
    A = np.zeros((n, n), dtype=np.complex128)
    idx = np.arange(n)
    A[idx, idx] = 4.0 - k2 + 0.01j * freq
    A[idx[1:], idx[:-1]] = -1.0
    A[idx[:-1], idx[1:]] = -1.0
    return A


def test_eigensolver_on_hessenberg_matches_lapack():
    rng = np.random.default_rng(1)
    for n in (5, 17, 40):
        H = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), -1)
        vals, vecs = G.hessenberg_qr_eigen(H)
        ref = np.linalg.eigvals(H)
        for v in vals:
            assert np.min(np.abs(ref - v)) <= 1e-9 * np.abs(ref).max()
        for v in ref:
            assert np.min(np.abs(vals - v)) <= 1e-9 * np.abs(ref).max()
        for j in range(n):
            y = vecs[:, j]
            assert np.linalg.norm(H @ y - vals[j] * y) <= 1e-8 * np.linalg.norm(H) * np.linalg.norm(y)


def test_dense_to_hessenberg_similarity_and_the_complex_quirk():
    rng = np.random.default_rng(2)
    n = 12
    Ar = rng.standard_normal((n, n)).astype(np.complex128)
    h, q = G.dense_to_hessenberg(Ar)
    assert np.allclose(q.conj().T @ q, np.eye(n), atol=1e-13)
    full = q.conj().T @ Ar @ q
    assert np.allclose(np.triu(full, -1), h, atol=1e-12)
    assert np.abs(np.tril(full, -2)).max() < 1e-12  # real data: an exact reduction
    ref = np.linalg.eigvals(Ar)
    assert all(np.min(np.abs(ref - v)) < 1e-10 for v in np.linalg.eigvals(h))
    Ac = Ar + 1j * rng.standard_normal((n, n))
    h, q = G.dense_to_hessenberg(Ac)
    assert np.allclose(q.conj().T @ q, np.eye(n), atol=1e-13)
    full = q.conj().T @ Ac @ q
    assert np.allclose(np.triu(full, -1), h, atol=1e-12)
    # complex data: the real-alpha reflector leaves weight below the subdiagonal, which is dropped
    assert np.abs(np.tril(full, -2)).max() > 1e-3


def test_first_solve_seeds_the_subspace():
    n = 40
    A = helmholtz_like(n, 1.0)
    b = np.sin(np.arange(n)) + 0j
    sub = G.Subspace(8)
    x, info = G.gcro_dr(A, b, sub, restart=30, max_iterations=100, tol=1e-10)
    err = np.linalg.norm(A @ x - b)
    assert info["converged"] and abs(info["residual"] - err / np.linalg.norm(b)) < 1e-12 and err < 1e-7
    assert 0 < len(sub.u) <= 8
    U = np.stack(sub.u)
    assert np.allclose(U.conj() @ U.T, np.eye(len(sub.u)), atol=1e-10)


def test_recycling_does_not_slow_a_nearby_solve_and_matches_lapack():
    n = 60
    b = np.cos(np.arange(n)) + 0.5j
    sub = G.Subspace(10)
    _, i0 = G.gcro_dr(helmholtz_like(n, 1.0), b, sub, restart=30, tol=1e-9)
    A1 = helmholtz_like(n, 1.05)
    x1, i1 = G.gcro_dr(A1, b, sub, restart=30, tol=1e-9)
    assert i0["converged"] and i1["converged"] and i1["recycled_in"] > 0
    assert i1["iterations"] <= i0["iterations"]
    assert np.linalg.norm(x1 - np.linalg.solve(A1, b)) <= 1e-7 * np.linalg.norm(x1)
    A = helmholtz_like(32, 0.5)
    b = np.arange(32) * 0.1 + 0j
    x, info = G.gcro_dr(A, b, G.Subspace(6), restart=30, tol=1e-12)
    assert info["converged"] and np.linalg.norm(x - np.linalg.solve(A, b)) < 1e-7


def test_sweep_recycling_at_least_halves_the_arnoldi_steps():
    """frequency_sweep_recycling_beats_independent_solves (gcro_dr.rs:1222-1300),
    60 of its 100 frequencies; the plain total counts the same GCRO-DR with an
    empty subspace, which is GMRES(40)."""
    n, b = 800, np.cos(np.arange(800)) + 0.5j
    sub = G.Subspace(20)
    indep = recycled = 0
    for f_idx in range(60):
        A = helmholtz_like(n, 1.0 + 0.05 * f_idx)
        _, plain = G.gcro_dr(A, b, G.Subspace(0), restart=40, max_iterations=200, tol=1e-8)
        x, rec = G.gcro_dr(A, b, sub, restart=40, max_iterations=200, tol=1e-8)
        assert plain["converged"] and rec["converged"]
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-4
        indep += plain["iterations"]
        recycled += rec["iterations"]
    print(f"sweep: independent {indep}, recycled {recycled}, ratio {indep / recycled:.2f}")
    assert indep / recycled >= 2.0


def test_diagonal_smoke_and_gated_restart():
    A = np.diag([2.0, 3.0]).astype(np.complex128)
    x, info = G.gcro_dr(A, np.array([2.0, 6.0], complex), G.Subspace(2), restart=5, max_iterations=10, tol=1e-12)
    assert info["converged"] and abs(x[0] - 1) < 1e-9 and abs(x[1] - 2) < 1e-9
    # rejected_exact_candidate_restarts_gcro_dr (gcro_dr.rs:1351-1387)
    M = lambda r: np.array([r[0], 1e-4 * r[1]])
    x, info = G.gcro_dr(np.eye(2, dtype=complex), np.array([1.0, 1.0], complex), G.Subspace(0), restart=2,
                        max_iterations=2, tol=1e-3, precond=M)
    assert info["converged"] and info["iterations"] >= 2


def test_failed_solve_leaves_the_subspace_unchanged():
    n = 30
    A = helmholtz_like(n, 0.7)
    b = np.ones(n, complex)
    sub = G.Subspace(5)
    G.gcro_dr(A, b, sub, restart=20, tol=1e-10)
    before = [u.copy() for u in sub.u]
    calls = {"n": 0}

    def bad(v):
        calls["n"] += 1
        return A @ v if calls["n"] < 12 else np.full(n, np.nan, complex)

    with pytest.raises(FloatingPointError):
        G.gcro_dr(bad, b, sub, restart=20, tol=1e-10)
    assert len(sub.u) == len(before) and all(np.array_equal(u, v) for u, v in zip(sub.u, before))
    with pytest.raises(ValueError):
        G.gcro_dr(A, np.full(n, np.inf, complex), sub)
