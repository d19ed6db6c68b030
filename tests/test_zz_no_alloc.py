"""The allocation-free block product (sbh_operator_apply_block_noalloc).

It runs the same kernels as sbh_operator_apply_block with its split-K
partial sums in the library's fixed per-GPU scratch instead of the context's
growable one, so on the same device inputs both entry points give the same
bits.  Two calls in a row, the second on a smaller operator and block, check
that nothing left in the shared scratch by the first reaches the second; the
products are also checked against numpy on the downloaded matrix.
"""

import numpy as np
import pytest

import sonium_amd as sa
from sonium_amd.mesh import icosphere

pytestmark = pytest.mark.gpu


def _system(ctx, level, k):
    m = icosphere(1.0, level)
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    system = sa.build_tbem_system_with_beta(ctx, tm, ph, ph.burton_miller_beta())
    return tm, system


def _block(n, k, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))


def _both(ctx, op, X):
    n, k = X.shape
    dx = sa.DeviceVector(ctx, n * k)
    dx.upload(X)
    y_ref = sa.DeviceVector(ctx, n * k)
    y_na = sa.DeviceVector(ctx, n * k)
    op.apply_block_device(dx, y_ref, k)
    op.apply_block_noalloc(dx, y_na, k)
    Y_ref = y_ref.download().reshape(n, k)
    Y_na = y_na.download().reshape(n, k)
    for v in (dx, y_ref, y_na):
        v.close()
    return Y_ref, Y_na


def test_noalloc_block_product_matches_apply_block(ctx):
    tm1, sys1 = _system(ctx, 3, 2.0)   # 1280 panels
    tm2, sys2 = _system(ctx, 2, 1.5)   # 320 panels
    op1 = sa.DenseMatrixOperator(sys1)
    op2 = sa.DenseMatrixOperator(sys2)
    A1 = sys1.matrix
    A2 = sys2.matrix
    cases = [(op1, A1, 64, 1), (op2, A2, 5, 2), (op1, A1, 1, 3), (op2, A2, 64, 4), (op2, A2, 3, 5)]
    for op, A, k, seed in cases:
        X = _block(A.shape[0], k, seed)
        Y_ref, Y_na = _both(ctx, op, X)
        ref = A @ X
        err = np.linalg.norm(Y_na - ref) / np.linalg.norm(ref)
        print(f"n={A.shape[0]} k={k}: identical bits {np.array_equal(Y_na, Y_ref)}, vs numpy {err:.2e}")
        assert np.array_equal(Y_na, Y_ref)
        assert err <= 1e-13
    for o in (op1, op2, sys1, sys2, tm1, tm2):
        o.close()


def test_noalloc_block_product_rejects_what_does_not_fit(ctx):
    # 8 splits x n x 128 partial columns exceed the 64 MiB scratch for n > 8192
    # at 64 columns; the shape is refused on the host before any launch.
    n = 8320
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    X = sa.DeviceVector(ctx, n * 64)
    Y = sa.DeviceVector(ctx, n * 64)
    with pytest.raises(sa.SbhError) as e:
        op.apply_block_noalloc(X, Y, 64)
    assert e.value.status == 1 and "MiB" in str(e.value)
    for o in (X, Y, op, system):
        o.close()
