"""Shared-basis block GMRES on the device (sbh_block_gmres) against the
oracle's restatement of block_gmres_inner (block_gmres.rs:666-1120) and LAPACK.

The device and the oracle run the same recurrence (block Arnoldi on one
basis, block CGS with the DGKS trigger, per-column CGS2 QR with rank-drop
deflation, Givens least squares after each block step) with different
summation orders, so they take the same number of block steps and their
solutions agree far below the 1e-8 tolerance.  Cases: the 1280-panel sphere
with 12 and 64 plane waves, restarts, the Jacobi preconditioner, an initial
guess, duplicated and zero right-hand sides (deflation), exact-residual
checkpoints, one right-hand side (scalar GMRES), device-resident blocks, and
the checked boundary's rejections.
"""

import numpy as np
import pytest

import oracle
from oracle import mie
import sonium_amd as sa
from sonium_amd.mesh import fibonacci_directions, icosphere

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sphere1280(ctx):
    m = icosphere(1.0, 3)
    k = 2.0
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    system = sa.build_tbem_system_with_beta(ctx, tm, ph, beta)
    A = system.matrix
    c, nn, area = m.element_geometry()
    dirs = fibonacci_directions(64)
    B = np.stack([oracle.plane_wave_rhs(c, nn, k, beta, direction=d) for d in dirs], 1)
    yield {"mesh": m, "tm": tm, "sys": system, "A": A, "B": B, "dirs": dirs, "k": k, "area": area, "c": c}
    system.close()
    tm.close()


def _compare(A, B, sol, ref, st, label, steps_slack=1):
    X = sol.x
    Xl = np.linalg.solve(A, B)
    err_o = np.linalg.norm(X - ref, axis=0) / np.linalg.norm(ref, axis=0)
    err_l = np.linalg.norm(X - Xl, axis=0) / np.linalg.norm(Xl, axis=0)
    res = np.linalg.norm(B - A @ X, axis=0) / np.linalg.norm(B, axis=0)
    print(f"{label}: {sol.iterations} block steps (oracle {st['iterations']}), basis {sol.max_basis}, "
          f"max true residual {res.max():.2e} (reported {sol.true_residuals.max():.2e}), "
          f"vs oracle {err_o.max():.2e}, vs LAPACK {err_l.max():.2e}")
    assert abs(sol.iterations - st["iterations"]) <= steps_slack
    assert sol.converged == st["converged"]
    np.testing.assert_allclose(sol.true_residuals, res, rtol=1e-6, atol=1e-14)
    np.testing.assert_allclose(sol.true_residuals, st["true_residuals"], rtol=0.5, atol=2e-10)
    return err_o, err_l, res


@pytest.mark.parametrize("s", [12, 64])
def test_block_gmres_matches_oracle_and_lapack(sphere1280, ctx, s):
    A, B = sphere1280["A"], sphere1280["B"][:, :s]
    op = sa.DenseMatrixOperator(sphere1280["sys"])
    cfg = sa.BlockGmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    sol = sa.block_gmres(op, B, cfg)
    ref, st = oracle.block_gmres(A, B, restart=50, max_iterations=100, tol=1e-8)
    err_o, err_l, res = _compare(A, B, sol, ref, st, f"{s} plane waves")
    assert sol.converged and res.max() <= 1e-8
    assert err_o.max() <= 1e-7 and err_l.max() <= 1e-7
    # the shared basis: fewer block steps than the single-RHS solve needs
    x1, st1 = oracle.gmres(A, B[:, 0], restart=50, max_iterations=100, tol=1e-8)
    assert sol.iterations < st1["iterations"]
    # surface pressure against the Mie series per direction
    m = sphere1280
    for col in (0, s - 1):
        p = mie.surface_pressure(m["k"], 1.0, m["c"], direction=m["dirs"][col])
        assert mie.weighted_relative_l2(sol.x[:, col], p, m["area"]) < 0.02
    op.close()


def test_block_gmres_restarts_jacobi_and_guess(sphere1280):
    A, B = sphere1280["A"], sphere1280["B"][:, :12]
    op = sa.DenseMatrixOperator(sphere1280["sys"])
    # restart 3: several cycles
    cfg = sa.BlockGmresConfig(max_iterations=100, restart=3, tolerance=1e-8)
    sol = sa.block_gmres(op, B, cfg)
    ref, st = oracle.block_gmres(A, B, restart=3, max_iterations=100, tol=1e-8)
    _, err_l, res = _compare(A, B, sol, ref, st, "restart 3")
    assert sol.restart_cycles > 1 and res.max() <= 1e-8 and err_l.max() <= 1e-7
    # Jacobi
    pc = sa.JacobiPreconditioner(op)
    cfg = sa.BlockGmresConfig(max_iterations=100, restart=30, tolerance=1e-8)
    sol = sa.block_gmres_preconditioned_with_guess(op, B, pc, None, cfg)
    ref, st = oracle.block_gmres(A, B, restart=30, max_iterations=100, tol=1e-8, jacobi=True)
    _, err_l, res = _compare(A, B, sol, ref, st, "Jacobi")
    assert res.max() <= 1e-8 and err_l.max() <= 1e-7
    # initial guess: the LAPACK solution perturbed
    Xl = np.linalg.solve(A, B)
    rng = np.random.default_rng(3)
    X0 = Xl + 1e-3 * (rng.standard_normal(Xl.shape) + 1j * rng.standard_normal(Xl.shape))
    sol = sa.block_gmres_preconditioned_with_guess(op, B, None, X0, cfg)
    ref, st = oracle.block_gmres(A, B, X0=X0, restart=30, max_iterations=100, tol=1e-8)
    _, err_l, res = _compare(A, B, sol, ref, st, "initial guess")
    assert res.max() <= 1e-8
    # exact guess: converged at the first residual, no block step
    sol = sa.block_gmres_preconditioned_with_guess(op, B, None, Xl, cfg)
    assert sol.converged and sol.iterations == 0
    op.close()


def test_block_gmres_deflation_and_zero_columns(sphere1280):
    A = sphere1280["A"]
    B0 = sphere1280["B"][:, :4]
    # duplicated, scaled and zero columns (block_gmres.rs:1383-1500 cases)
    B = np.concatenate([B0, B0[:, :2], 3.0 * B0[:, 1:2], np.zeros((A.shape[0], 1))], axis=1)
    op = sa.DenseMatrixOperator(sphere1280["sys"])
    cfg = sa.BlockGmresConfig(max_iterations=50, restart=30, tolerance=1e-8)
    sol = sa.block_gmres(op, B, cfg)
    ref, st = oracle.block_gmres(A, B, restart=30, max_iterations=50, tol=1e-8)
    live = slice(0, B.shape[1] - 1)
    X = sol.x
    res = np.linalg.norm(B[:, live] - A @ X[:, live], axis=0) / np.linalg.norm(B[:, live], axis=0)
    err_o = np.linalg.norm(X[:, live] - ref[:, live], axis=0) / np.linalg.norm(ref[:, live], axis=0)
    print(f"deflation: {sol.iterations} block steps (oracle {st['iterations']}), {sol.deflated} directions "
          f"dropped, max true residual {res.max():.2e}, vs oracle {err_o.max():.2e}")
    assert abs(sol.iterations - st["iterations"]) <= 1 and sol.converged == st["converged"]
    assert err_o.max() <= 1e-7
    # the initial block QR drops both duplicates, the multiple and the zero column
    assert sol.deflated >= 4
    assert np.all(sol.x[:, -1] == 0) and sol.true_residuals[-1] == 0.0
    np.testing.assert_allclose(sol.x[:, 4:6], sol.x[:, :2], rtol=0, atol=1e-9 * np.abs(sol.x).max())
    np.testing.assert_allclose(sol.x[:, 6], 3.0 * sol.x[:, 1], rtol=0, atol=1e-8 * np.abs(sol.x).max())
    assert res.max() <= 1e-8
    op.close()


def test_block_gmres_checkpoints_single_rhs_and_device(sphere1280, ctx):
    A, B = sphere1280["A"], sphere1280["B"][:, :8]
    op = sa.DenseMatrixOperator(sphere1280["sys"])
    cfg = sa.BlockGmresConfig(max_iterations=100, restart=40, tolerance=1e-8, exact_residual_check_interval=3)
    sol = sa.block_gmres(op, B, cfg)
    ref, st = oracle.block_gmres(A, B, restart=40, max_iterations=100, tol=1e-8, exact_residual_check_interval=3)
    _compare(A, B, sol, ref, st, "exact checkpoints every 3 steps")
    assert sol.exact_residual_checks >= sol.iterations // 3
    # one right-hand side: the block recurrence is scalar GMRES
    b = B[:, :1]
    sol1 = sa.block_gmres(op, b, sa.BlockGmresConfig(max_iterations=100, restart=50, tolerance=1e-8))
    x_s, st_s = oracle.gmres(A, b[:, 0], restart=50, max_iterations=100, tol=1e-8)
    print(f"one right-hand side: {sol1.iterations} block steps, scalar GMRES {st_s['iterations']}")
    assert abs(sol1.iterations - st_s["iterations"]) <= 1
    assert np.linalg.norm(sol1.x[:, 0] - x_s) / np.linalg.norm(x_s) <= 1e-7
    # device-resident blocks give the host-path solution
    n, s = B.shape
    dB = sa.DeviceVector(ctx, n * s)
    dB.upload(B)
    dX = sa.DeviceVector(ctx, n * s)
    cfg = sa.BlockGmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    sd = sa.block_gmres_device(op, dB, s, dX, cfg)
    sh = sa.block_gmres(op, B, cfg)
    Xd = dX.download().reshape(n, s)
    assert sd.iterations == sh.iterations and np.array_equal(Xd, sh.x)
    dB.close()
    dX.close()
    op.close()


def test_block_gmres_checked_boundary(sphere1280):
    A, B = sphere1280["A"], sphere1280["B"][:, :2].copy()
    op = sa.DenseMatrixOperator(sphere1280["sys"])
    for bad in (sa.BlockGmresConfig(max_iterations=0), sa.BlockGmresConfig(restart=0),
                sa.BlockGmresConfig(tolerance=0.0), sa.BlockGmresConfig(tolerance=float("nan"))):
        with pytest.raises(sa.SbhError) as e:
            sa.block_gmres(op, B, bad)
        assert e.value.status == 1
    Bn = B.copy()
    Bn[3, 1] = np.nan
    with pytest.raises(sa.SbhError) as e:
        sa.block_gmres(op, Bn)
    assert "right-hand side" in str(e.value)
    X0 = np.zeros_like(B)
    X0[0, 0] = np.inf
    with pytest.raises(sa.SbhError) as e:
        sa.block_gmres_preconditioned_with_guess(op, B, None, X0)
    assert "initial guess" in str(e.value)
    op.close()
