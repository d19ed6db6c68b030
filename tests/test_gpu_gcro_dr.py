"""GCRO-DR on the device (sbh_gcro_dr) against the oracle's restatement of
gcro_dr_core (oracle/gcro_dr.py, gcro_dr.rs:292-1050).

Both run the same recurrence (recycled C-projection, CGS+DGKS Arnoldi,
Givens least squares, harmonic Ritz refresh with the reference's small
eigensolver) with different summation orders, so each solve of a sequence
takes the same number of Arnoldi steps and their solutions agree far below
the tolerance.  Cases: the reference's tridiagonal Helmholtz-like sweep as a
caller-supplied operator, a 1280-panel Burton-Miller sweep over six
wavenumbers with the dense operator, device-resident vectors, and the checked
boundary (invalid configuration, non-finite input, a failing operator leaves
the subspace unchanged).
"""

import numpy as np
import pytest

from oracle import gcro_dr as og
import oracle
import sonium_amd as sa
from sonium_amd.mesh import icosphere

pytestmark = pytest.mark.gpu


def helmholtz_like(n, freq):
    k2 = freq * freq * 0.05
    A = np.zeros((n, n), dtype=np.complex128)
    i = np.arange(n)
    A[i, i] = 4.0 - k2 + 0.01j * freq
    A[i[1:], i[:-1]] = -1.0
    A[i[:-1], i[1:]] = -1.0
    return A


def _principal_cos(U, W):
    """Smallest cosine of the principal angles between the row spaces of U and W."""
    qu = np.linalg.qr(U.T)[0]
    qw = np.linalg.qr(W.T)[0]
    return np.linalg.svd(qu.conj().T @ qw, compute_uv=False).min()


def test_sweep_with_callback_operator_matches_oracle(ctx):
    n = 120
    b = np.cos(np.arange(n)) + 0.5j
    sub = sa.GcroDrSubspace(ctx, n, 10)
    osub = og.Subspace(10)
    cfg = sa.GcroDrConfig(max_iterations=50, restart=30, tolerance=1e-9)
    steps_dev, steps_or = [], []
    for f_idx in range(8):
        A = helmholtz_like(n, 1.0 + 0.07 * f_idx)
        op = sa.CallbackOperator(ctx, n, lambda v, A=A: A @ v)
        sol = sa.gcro_dr(op, b, None, sub, cfg)
        xo, info = og.gcro_dr(A, b, osub, restart=30, max_iterations=50, tol=1e-9)
        steps_dev.append(sol.iterations)
        steps_or.append(info["iterations"])
        assert sol.converged and info["converged"]
        assert abs(sol.iterations - info["iterations"]) <= 1
        assert np.linalg.norm(sol.x - xo) <= 1e-8 * np.linalg.norm(xo)
        assert sol.residual <= 1e-9
        assert sub.num_vectors() == len(osub.u)
        if osub.u:
            assert _principal_cos(sub.vectors(), np.stack(osub.u)) > 1 - 1e-6
        op.close()
    print("Arnoldi steps per solve, device", steps_dev, "oracle", steps_or)
    assert sum(steps_dev[1:]) < 7 * steps_dev[0]
    sub.close()


@pytest.fixture(scope="module")
def sphere_sweep(ctx):
    m = icosphere(1.0, 3)
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    c, nn, _ = tm.geometry()
    yield {"mesh": m, "tm": tm, "ws": ws, "c": c, "nn": nn, "n": m.num_elements}
    ws.close()
    tm.close()


def test_burton_miller_sweep_matches_oracle(sphere_sweep, ctx):
    s = sphere_sweep
    n = s["n"]
    sub = sa.GcroDrSubspace(ctx, n, 10)
    osub = og.Subspace(10)
    cfg = sa.GcroDrConfig(max_iterations=20, restart=50, tolerance=1e-8)
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    plain, rec = [], []
    for k in np.linspace(2.0, 2.25, 6):
        ph = sa.PhysicsParams.from_wavenumber(float(k))
        beta = ph.burton_miller_beta()
        s["ws"].assemble(ph, beta, system)
        A = system.matrix
        b = oracle.plane_wave_rhs(s["c"], s["nn"], float(k), beta)
        sol = sa.gcro_dr(op, b, None, sub, cfg)
        xo, info = og.gcro_dr(A, b, osub, restart=50, max_iterations=20, tol=1e-8)
        g = sa.gmres(op, b, sa.GmresConfig(max_iterations=20, restart=50, tolerance=1e-8))
        plain.append(g.iterations)
        rec.append(sol.iterations)
        assert sol.converged and info["converged"]
        assert abs(sol.iterations - info["iterations"]) <= 1
        assert np.linalg.norm(sol.x - xo) <= 1e-7 * np.linalg.norm(xo)
        res = np.linalg.norm(b - A @ sol.x) / np.linalg.norm(b)
        assert res <= 1e-8 and abs(res - sol.residual) <= 1e-12
    print("1280 panels, k 2.0..2.25: GMRES steps", plain, "GCRO-DR steps", rec,
          "products incl. the rebuild of C", [r + 10 for r in rec[1:]])
    assert sum(rec[1:]) < sum(plain[1:])
    op.close()
    system.close()
    sub.close()


def test_device_vectors_and_checked_boundary(sphere_sweep, ctx):
    s = sphere_sweep
    n = s["n"]
    ph = sa.PhysicsParams.from_wavenumber(2.0)
    beta = ph.burton_miller_beta()
    system = s["ws"].assemble(ph, beta)
    op = sa.DenseMatrixOperator(system)
    b = oracle.plane_wave_rhs(s["c"], s["nn"], 2.0, beta)
    cfg = sa.GcroDrConfig(max_iterations=20, restart=40, tolerance=1e-8)
    # device-resident b and x give the host-path result
    sub_h, sub_d = sa.GcroDrSubspace(ctx, n, 6), sa.GcroDrSubspace(ctx, n, 6)
    sh = sa.gcro_dr(op, b, None, sub_h, cfg)
    db, dx = sa.DeviceVector(ctx, n), sa.DeviceVector(ctx, n)
    db.upload(b)
    sd = sa.gcro_dr_device(op, db.ptr, dx, sub_d, cfg)
    assert sd.iterations == sh.iterations and np.array_equal(dx.download(), sh.x)
    assert np.array_equal(sub_d.vectors(), sub_h.vectors())
    # checked boundary
    for bad in (sa.GcroDrConfig(max_iterations=0), sa.GcroDrConfig(restart=0),
                sa.GcroDrConfig(tolerance=float("nan"))):
        with pytest.raises(sa.SbhError) as e:
            sa.gcro_dr(op, b, None, sub_h, bad)
        assert e.value.status == 1
    bn = b.copy()
    bn[5] = np.nan
    with pytest.raises(sa.SbhError) as e:
        sa.gcro_dr(op, bn, None, sub_h, cfg)
    assert "right-hand side" in str(e.value)
    # a failing operator: the call errors and the recycled subspace is unchanged
    before = sub_h.vectors()
    A = system.matrix
    calls = {"n": 0}

    def flaky(v):
        calls["n"] += 1
        return A @ v if calls["n"] < 9 else np.full(n, np.nan, complex)

    fop = sa.CallbackOperator(ctx, n, flaky)
    with pytest.raises(sa.SbhError):
        sa.gcro_dr(fop, b, None, sub_h, cfg)
    assert np.array_equal(sub_h.vectors(), before)
    # reset empties it; k = 0 never recycles
    sub_h.reset()
    assert sub_h.num_vectors() == 0 and not sub_h.is_populated()
    sub0 = sa.GcroDrSubspace(ctx, n, 0)
    s0 = sa.gcro_dr(op, b, None, sub0, cfg)
    assert s0.converged and sub0.num_vectors() == 0
    for o in (fop, sub0, sub_h, sub_d, db, dx, op, system):
        o.close()
