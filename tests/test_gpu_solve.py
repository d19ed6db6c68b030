"""GPU parity of the operator, preconditioner and GMRES against the oracle and Mie.

* apply / apply_transpose / apply_hermitian / apply_block against the same
  products computed with numpy on the downloaded matrix (the reference's
  tbem.rs:6627-6705 checks forward, transpose and Hermitian products);
* GMRES (gmres.rs:631-1057) against the oracle's restatement on the same
  system: solutions agree to 1e-8 relative, the recomputed residual meets
  the tolerance;
* end to end against the rigid-sphere Mie series (solutions_3d.rs:315-387).
"""

import numpy as np
import pytest

import oracle
from oracle import mie
import sonium_amd as sa
from sonium_amd.mesh import icosphere

from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1(ctx):
    """Config C1: icosphere s=3 (1280 panels), k=2, beta=i/k, plane wave +z."""
    m = icosphere(1.0, 3)
    k = 2.0
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    sys = ws.assemble(ph, beta)
    sa.IncidentField.plane_wave_z().add_to_system(sys, tm, ph, beta, sign=1.0)
    return {"mesh": m, "tm": tm, "ph": ph, "beta": beta, "sys": sys, "ws": ws,
            "A": sys.matrix, "b": sys.rhs}


def test_rhs_includes_incident_field(c1):
    m = c1["mesh"]
    c, n, _ = m.element_geometry()
    ref = oracle.plane_wave_rhs(c, n, 2.0, c1["beta"])
    assert np.max(np.abs(c1["b"] - ref)) <= 1e-14


def test_operator_products(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    A = c1["A"]
    rng = np.random.default_rng(11)
    n = A.shape[0]
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert op.num_rows() == n and op.num_cols() == n
    assert op.transpose_capability() == "Exact" and op.hermitian_capability() == "Exact"
    assert rel_l2(op.apply(x), A @ x) < 1e-14
    assert rel_l2(op.apply_transpose(x), A.T @ x) < 1e-14
    assert rel_l2(op.apply_hermitian(x), A.conj().T @ x) < 1e-14
    X = rng.standard_normal((n, 5)) + 1j * rng.standard_normal((n, 5))
    assert rel_l2(op.apply_block(X), A @ X) < 1e-14
    # checked apply rejects non-finite input before any device work
    bad = x.copy()
    bad[3] = np.nan
    with pytest.raises(sa.SbhError):
        op.apply(bad)


def test_matvec_ragged_sizes(ctx):
    """Row counts that do not fill a wave's 4-row block; columns not a multiple of 256."""
    for subdiv, nm in ((0, None), (1, None)):
        m = icosphere(1.0, subdiv)
        tm = sa.TbemMesh(ctx, m)
        ph = sa.PhysicsParams.from_wavenumber(1.5)
        sys = sa.build_tbem_system(ctx, tm, ph)
        A = sys.matrix
        op = sa.DenseMatrixOperator(sys)
        x = np.arange(A.shape[0]) * (0.3 + 0.1j)
        assert rel_l2(op.apply(x), A @ x) < 1e-14


def test_gmres_matches_oracle_and_lu(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    cfg = sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    sol = sa.gmres(op, c1["b"], cfg)
    x_ref, st = oracle.gmres(c1["A"], c1["b"], restart=50, max_iterations=100, tol=1e-8)
    x_lu = np.linalg.solve(c1["A"], c1["b"])
    print("gpu", sol.iterations, sol.residual, "cpu", st["iterations"], st["residual"])
    assert sol.converged and sol.status == "converged"
    assert sol.residual <= 1e-8
    assert sol.iterations == st["iterations"]
    assert rel_l2(sol.x, x_ref) <= 1e-8
    assert rel_l2(sol.x, x_lu) <= 1e-7


def test_gmres_surface_pressure_vs_mie(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    sol = sa.gmres(op, c1["b"], sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8))
    c, _, area = c1["mesh"].element_geometry()
    p_mie = mie.surface_pressure(2.0, 1.0, c)
    err = mie.weighted_relative_l2(sol.x, p_mie, area)
    print("surface L2 vs Mie", err)
    # beta = i/k on 1280 panels: the oracle gives 0.57 %; the same discretisation
    # must give the same error to many digits.
    assert err < 0.01


def test_gmres_jacobi_and_guess(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    M = sa.JacobiPreconditioner(op)
    cfg = sa.GmresConfig(max_iterations=100, restart=30, tolerance=1e-10)
    sol = sa.gmres_preconditioned(op, M, c1["b"], cfg)
    x_ref, st = oracle.gmres(c1["A"], c1["b"], restart=30, max_iterations=100, tol=1e-10, jacobi=True)
    assert sol.converged and st["converged"]
    assert rel_l2(sol.x, x_ref) <= 1e-8
    # a good initial guess converges immediately (gmres.rs:700-715 early return)
    sol2 = sa.gmres_with_guess(op, c1["b"], sol.x, cfg)
    assert sol2.converged and sol2.iterations == 0


def test_gmres_rejects_invalid_inputs(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    b = c1["b"].copy()
    b[0] = np.inf
    sol = sa.gmres(op, b, sa.GmresConfig(max_iterations=10, restart=10, tolerance=1e-8))
    assert sol.status == "numerical_error" and not sol.converged and sol.iterations == 0
    sol = sa.gmres(op, c1["b"], sa.GmresConfig(max_iterations=0, restart=10, tolerance=1e-8))
    assert sol.status == "numerical_error"


def test_gmres_zero_rhs(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    sol = sa.gmres(op, np.zeros(op.num_rows(), dtype=complex), sa.GmresConfig(10, 10, 1e-8))
    x_ref, st = oracle.gmres(c1["A"], np.zeros(op.num_rows(), dtype=complex), restart=10, max_iterations=10, tol=1e-8)
    assert sol.converged == st["converged"]
    assert np.all(sol.x == 0)


def test_gmres_restart_cycles_match_oracle(c1):
    """Short restart forces several cycles and exercises the restart bookkeeping."""
    op = sa.DenseMatrixOperator(c1["sys"])
    cfg = sa.GmresConfig(max_iterations=50, restart=8, tolerance=1e-9)
    sol = sa.gmres(op, c1["b"], cfg)
    x_ref, st = oracle.gmres(c1["A"], c1["b"], restart=8, max_iterations=50, tol=1e-9)
    assert sol.converged == st["converged"]
    assert sol.restarts == st["restarts"]
    assert abs(sol.iterations - st["iterations"]) <= 1
    assert rel_l2(sol.x, x_ref) <= 1e-7


@pytest.mark.parametrize("interval", [0, 4])
def test_gmres_convergence_check_interval_matches_oracle(c1, interval):
    """GmresConfig::convergence_check_interval (gmres.rs:871-876)."""
    op = sa.DenseMatrixOperator(c1["sys"])
    cfg = sa.GmresConfig(max_iterations=10, restart=50, tolerance=1e-8, convergence_check_interval=interval,
                         exact_residual_check_interval=3)
    sol = sa.gmres(op, c1["b"], cfg)
    x_ref, st = oracle.gmres(c1["A"], c1["b"], restart=50, max_iterations=10, tol=1e-8,
                             convergence_check_interval=interval)
    assert sol.converged and st["converged"]
    assert sol.iterations == st["iterations"] and sol.restarts == st["restarts"]
    assert rel_l2(sol.x, x_ref) <= 1e-8
    # every converged return is confirmed by one recomputed residual whatever the interval
    assert sol.exact_residual_checks == 1


def test_gmres_adaptive_restart_matches_oracle(c1):
    """min_restart / max_restart (gmres.rs:1013-1029): a short restart on C1
    stalls and grows; the cycle lengths, and so the step count, match."""
    op = sa.DenseMatrixOperator(c1["sys"])
    cfg = sa.GmresConfig(max_iterations=30, restart=4, tolerance=1e-9, min_restart=2, max_restart=40)
    sol = sa.gmres(op, c1["b"], cfg)
    x_ref, st = oracle.gmres(c1["A"], c1["b"], restart=4, max_iterations=30, tol=1e-9, min_restart=2,
                             max_restart=40)
    _, fixed = oracle.gmres(c1["A"], c1["b"], restart=4, max_iterations=30, tol=1e-9)
    print("adaptive", sol.iterations, sol.restarts, "oracle", st["iterations"], st["restarts"], "fixed",
          fixed["iterations"], fixed["restarts"], fixed["converged"])
    assert sol.converged == st["converged"]
    assert sol.restarts == st["restarts"] and abs(sol.iterations - st["iterations"]) <= 1
    assert st["iterations"] != fixed["iterations"]
    if st["converged"]:
        assert rel_l2(sol.x, x_ref) <= 1e-7


def test_gmres_rejects_negative_options(c1):
    op = sa.DenseMatrixOperator(c1["sys"])
    sol = sa.gmres(op, c1["b"], sa.GmresConfig(10, 10, 1e-8, convergence_check_interval=-1))
    assert sol.status == "numerical_error" and sol.iterations == 0
