"""Pins of the oracle's GMRES options (gmres.rs:26-46, 871-876, 1013-1029).

* convergence_check_interval: GMRES's projected residual never grows within
  a cycle, so checking every k-th step converges at the first checked step at
  or after the every-step solve's step (or the cycle's last step); interval 0
  behaves as 1 (the reference's regression test_gmres_convergence_check_interval_zero);
* adaptive restart: on the cyclic shift (the classic complete-stagnation
  example: GMRES(m) with m < n makes no progress from b = e_1) every cycle
  stalls, so the restart doubles each cycle up to max_restart; the iteration
  count has a closed form.  A run that halves is checked against the same
  cycles composed by hand from one-cycle solves (each fixed-restart cycle is
  pinned by LAPACK elsewhere), with the cycle length chosen by the ratio of
  true residuals.
"""

import numpy as np
import pytest

import oracle


def _system(n=60, seed=3):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4.0 + (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A.astype(np.complex128), b


@pytest.mark.parametrize("interval", [0, 2, 3, 5])
def test_convergence_check_interval(interval):
    A, b = _system()
    x1, s1 = oracle.gmres(A, b, restart=50, max_iterations=5, tol=1e-10)
    xk, sk = oracle.gmres(A, b, restart=50, max_iterations=5, tol=1e-10, convergence_check_interval=interval)
    assert s1["converged"] and sk["converged"] and s1["restarts"] == 0
    k = max(1, interval)
    j1 = s1["iterations"] - 1  # 0-based step of the every-step convergence
    expect = min(-(-j1 // k) * k, 50 - 1) + 1
    assert sk["iterations"] == expect
    assert np.linalg.norm(xk - np.linalg.solve(A, b)) / np.linalg.norm(xk) < 1e-9
    if k == 1:
        assert np.array_equal(x1, xk)


def test_adaptive_restart_doubles_on_stagnation():
    n = 64
    P = np.roll(np.eye(n), 1, axis=0).astype(np.complex128)  # cyclic shift: P e_i = e_{i+1}
    b = np.zeros(n, dtype=np.complex128)
    b[0] = 1.0
    _, fixed = oracle.gmres(P, b, restart=2, max_iterations=5, tol=1e-8)
    assert not fixed["converged"] and fixed["iterations"] == 10 and fixed["restarts"] == 5
    _, ad = oracle.gmres(P, b, restart=2, max_iterations=5, tol=1e-8, min_restart=2, max_restart=16)
    # cycles of 2, 4, 8, 16, 16 steps, none of which moves the residual
    assert not ad["converged"] and ad["iterations"] == 2 + 4 + 8 + 16 + 16
    assert ad["residual"] == pytest.approx(1.0, abs=1e-12)
    # with room for n steps the same problem is solved
    x, full = oracle.gmres(P, b, restart=2, max_iterations=8, tol=1e-8, min_restart=2, max_restart=64)
    assert full["converged"] and np.linalg.norm(P @ x - b) < 1e-8


def test_adaptive_restart_matches_composed_cycles():
    n = 120
    d = np.linspace(1.0, 30.0, n)
    rng = np.random.default_rng(5)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    A = (Q * d) @ Q.conj().T
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    m, mn, mx, cycles = 8, 2, 32, 6
    tol = 1e-13
    x = np.zeros(n, dtype=np.complex128)
    total, lengths = 0, []
    for _ in range(cycles):
        r0 = np.linalg.norm(b - A @ x)
        xn, st = oracle.gmres(A, b, x0=x, restart=m, max_iterations=1, tol=tol)
        total += st["iterations"]
        lengths.append(m)
        x = xn
        if st["converged"]:
            break
        ratio = np.linalg.norm(b - A @ x) / r0
        if ratio > 0.9 and m < mx:
            m = min(2 * m, mx)
        elif ratio < 0.1 and m > mn:
            m = max(m // 2, mn)
    xa, sa = oracle.gmres(A, b, restart=8, max_iterations=cycles, tol=tol, min_restart=mn, max_restart=mx)
    assert len(set(lengths)) > 1, "the composed run must change its restart"
    assert sa["iterations"] == total
    assert np.linalg.norm(xa - x) <= 1e-12 * np.linalg.norm(x)
