"""GPU tests of the matrix-free exact operator (apply_tbem_stabilized_exact,
hmatrix.rs:2363-2425).

The reference's own check (test_exact_stabilized_matvec_matches_dense_system,
hmatrix.rs:6785-6812) compares the exact product with the dense stabilized
matrix times x to 1e-11 per entry, on a mesh with an admittance panel and
beta_scale = 4.  Here:
* y = sign * A(beta) x agrees with the stored dense product of the same
  workspace to 1e-13 relative (summation order only), for rigid meshes and
  for meshes with pressure, admittance and nonzero-BC panels, sign = -1 and
  a scaled beta; and with the oracle's matrix to 1e-12;
* the block product equals the column-wise products;
* GMRES on the exact operator reproduces GMRES on the dense operator;
* products that need stored entries return SBH_ERR_UNSUPPORTED.
"""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import ellipsoid, icosphere

from helpers import rel_fro, rel_l2

pytestmark = pytest.mark.gpu


def _pair(ctx, m, k, beta, **bc):
    tm = sa.TbemMesh(ctx, m, **bc)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    ph = sa.PhysicsParams.from_wavenumber(k)
    system = ws.assemble(ph, beta)
    return tm, ws, ph, system


def test_exact_product_matches_dense_rigid(ctx):
    m = icosphere(1.0, 3)
    k = 2.0
    beta = 1j / k
    tm, ws, ph, system = _pair(ctx, m, k, beta)
    dense = sa.DenseMatrixOperator(system)
    exact = sa.TbemExactOperator(ws, ph, beta)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(m.num_elements) + 1j * rng.standard_normal(m.num_elements)
    y = exact.apply(x)
    assert rel_l2(y, dense.apply(x)) < 1e-13
    A_ref = oracle.build_tbem_system(m.nodes, m.triangles, k, beta)[0]
    assert rel_l2(y, A_ref @ x) < 1e-12
    assert exact.transpose_capability() == "Unsupported"
    assert dense.transpose_capability() == "Exact"


def test_exact_product_general_bc_sign_and_beta_scale(ctx):
    """Pressure, admittance and nonzero-BC panels; sign = -1; beta = 4 i/k (the
    reference test's beta_scale)."""
    m = ellipsoid([1.2, 0.8, 1.0], 2)
    n = m.num_elements
    rng = np.random.default_rng(3)
    k = 3.0
    beta = 4.0j / k
    bt = (rng.random(n) < 0.3).astype(np.int32)
    bv = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * (rng.random(n) < 0.5)
    adm = np.where(rng.random(n) < 0.3, 0.02 - 0.01j, np.nan)
    adm = np.where(bt == 0, adm, np.nan)
    tm, ws, ph, system = _pair(ctx, m, k, beta, bc_type=bt, bc_value=bv, admittance=adm)
    A = system.matrix
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = sa.TbemExactOperator(ws, ph, beta, sign=-1.0).apply(x)
    assert rel_l2(y, -(A @ x)) < 1e-13
    assert np.max(np.abs(y + A @ x)) < 1e-11  # the reference test's per-entry bound


def test_exact_block_product(ctx):
    m = icosphere(1.0, 2)
    k = 4.0
    tm, ws, ph, system = _pair(ctx, m, k, 1j / k)
    exact = sa.TbemExactOperator(ws, ph)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((m.num_elements, 3)) + 1j * rng.standard_normal((m.num_elements, 3))
    Y = exact.apply_block(X)
    for c in range(3):
        assert rel_l2(Y[:, c], exact.apply(X[:, c])) < 1e-15
    assert rel_fro(Y, system.matrix @ X) < 1e-13


def test_gmres_on_exact_operator(ctx):
    m = icosphere(1.0, 3)
    k = 2.0
    beta = 1j / k
    tm, ws, ph, system = _pair(ctx, m, k, beta)
    b = sa.IncidentField.plane_wave_z().compute_rhs_with_beta(tm, ph, beta)
    cfg = sa.GmresConfig(max_iterations=10, restart=50, tolerance=1e-8)
    d = sa.gmres(sa.DenseMatrixOperator(system), b, cfg)
    e = sa.gmres(sa.TbemExactOperator(ws, ph, beta), b, cfg)
    assert e.converged and e.residual < 1e-8
    assert abs(e.iterations - d.iterations) <= 1
    assert rel_l2(e.x, d.x) < 1e-8


def test_exact_operator_rejects_stored_entry_products(ctx):
    m = icosphere(1.0, 1)
    tm, ws, ph, system = _pair(ctx, m, 1.0, 1j)
    exact = sa.TbemExactOperator(ws, ph)
    x = np.ones(m.num_elements, dtype=np.complex128)
    for fn in (exact.apply_transpose, exact.apply_hermitian):
        with pytest.raises(sa.SbhError) as e:
            fn(x)
        assert e.value.status == 2
    with pytest.raises(sa.SbhError):
        sa.JacobiPreconditioner(exact)
    with pytest.raises(sa.SbhError):
        sa.gmres_block(exact, np.ones((m.num_elements, 2), dtype=np.complex128),
                       sa.GmresConfig(max_iterations=1, restart=5, tolerance=1e-8))
