"""Caller-supplied operators and preconditioners (the reference's gmres is
generic over LinearOperator and Preconditioner, gmres.rs:286,
traits.rs:2949-3290, 3645-3780).  A callback that multiplies by the assembled
matrix on the host reproduces the dense operator's solves (same step counts,
solutions to rounding), also in block GMRES and with a callback Jacobi
preconditioner; a failing or non-finite callback is a typed error."""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import fibonacci_directions, icosphere

pytestmark = pytest.mark.gpu


def test_callback_operator_and_preconditioner_match_the_library_ones(ctx):
    m = icosphere(1.0, 3)
    k = 2.0
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    system = sa.build_tbem_system_with_beta(ctx, tm, ph, beta)
    A = system.matrix
    n = A.shape[0]
    c, nn, _ = m.element_geometry()
    b = oracle.plane_wave_rhs(c, nn, k, beta)
    dense = sa.DenseMatrixOperator(system)
    calls = []

    def apply(x):
        calls.append(1)
        return A @ x

    cb = sa.CallbackOperator(ctx, n, apply)
    assert cb.num_rows() == n and cb.transpose_capability() == "Unsupported"
    x = np.random.default_rng(0).standard_normal(n) + 0j
    assert np.allclose(cb.apply(x), dense.apply(x), rtol=0, atol=1e-13 * np.abs(A @ x).max())
    cfg = sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    s_d = sa.gmres(dense, b, cfg)
    calls.clear()
    s_c = sa.gmres(cb, b, cfg)
    print(f"gmres: dense {s_d.iterations} steps, callback {s_c.iterations} steps ({len(calls)} callback products)")
    assert s_c.converged and s_c.iterations == s_d.iterations and s_c.residual <= 1e-8
    assert np.linalg.norm(s_c.x - s_d.x) / np.linalg.norm(s_d.x) <= 1e-10
    assert len(calls) == s_c.spmv_count
    # callback preconditioner = the library's Jacobi
    dinv = 1.0 / np.diag(A)
    pc = sa.CallbackPreconditioner(ctx, n, lambda r: dinv * r)
    pj = sa.JacobiPreconditioner(dense)
    s_j = sa.gmres_preconditioned_with_guess(dense, pj, b, None, cfg)
    s_p = sa.gmres_preconditioned_with_guess(cb, pc, b, None, cfg)
    assert s_p.converged and s_p.iterations == s_j.iterations
    assert np.linalg.norm(s_p.x - s_j.x) / np.linalg.norm(s_j.x) <= 1e-10
    # block GMRES over the callback operator (one callback per column)
    dirs = fibonacci_directions(6)
    B = np.stack([oracle.plane_wave_rhs(c, nn, k, beta, direction=d) for d in dirs], 1)
    bcfg = sa.BlockGmresConfig(max_iterations=50, restart=40, tolerance=1e-8)
    bd = sa.block_gmres(dense, B, bcfg)
    bc = sa.block_gmres(cb, B, bcfg)
    assert bc.converged and bc.iterations == bd.iterations
    assert np.max(np.linalg.norm(bc.x - bd.x, axis=0) / np.linalg.norm(bd.x, axis=0)) <= 1e-10
    assert np.allclose(cb.apply_block(B[:, :3]), A @ B[:, :3], rtol=0, atol=1e-12 * np.abs(A @ B[:, :3]).max())
    # failures: an exception and a non-finite result are typed errors; the context stays usable
    bad = sa.CallbackOperator(ctx, n, lambda x: 1 / 0)
    with pytest.raises(sa.SbhError) as e:
        sa.gmres(bad, b, cfg)
    assert e.value.status == 3 and "callback failed" in str(e.value)
    nan = sa.CallbackOperator(ctx, n, lambda x: x * np.nan)
    with pytest.raises(sa.SbhError) as e:
        nan.apply(x)
    assert e.value.status == 3
    with pytest.raises(sa.SbhError) as e:
        cb.apply_transpose(x)
    assert e.value.status == 2
    assert sa.gmres(dense, b, cfg).converged
    for o in (bad, nan, cb, pc, pj, dense, system, tm):
        o.close()
