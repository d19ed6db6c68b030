"""GPU tests of the batched path (many independent rigid scatterers).

* every instance matrix of the persistent batch assembly equals the
  single-system workspace assembly on the same mesh bit for bit (the same
  device functions), and the CPU oracle's matrix to 1e-12 relative
  Frobenius; the plane-wave right-hand sides agree with the single-system
  ones to 1e-15 and with the oracle to 1e-14;
* the lock-step FP64 batched GMRES reaches the recomputed relative residual
  1e-8 on every column, agrees with LAPACK on the oracle matrix to 1e-7 and
  takes the same number of Arnoldi steps as the single-system GMRES (within
  one, the reductions are ordered differently);
* FP16 and BF16 refinement reach the same FP64 residual;
* the graph-replayed cycle gives exactly the direct-launch result;
* the autotuner returns a valid shape and leaves results unchanged.
"""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import ellipsoid, uniform_directions

from helpers import rel_fro, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(ctx):
    rng = np.random.Generator(np.random.PCG64(7))
    specs = []
    for i, (level, k, nd) in enumerate([(1, 1.5, 1), (2, 3.0, 3), (1, 6.0, 5), (2, 0.7, 16), (2, 5.0, 2),
                                        (1, 2.2, 7)]):
        axes = rng.uniform(0.5, 1.5, size=3)
        specs.append({"mesh": ellipsoid(axes, level), "k": k, "directions": uniform_directions(rng, nd)})
    tms = [sa.TbemMesh(ctx, s["mesh"]) for s in specs]
    batch = sa.Batch(ctx, tms)
    batch.set_problem([s["k"] for s in specs], [s["directions"] for s in specs])
    batch.assemble()
    return {"specs": specs, "tms": tms, "batch": batch}


def test_batch_assembly_matches_single_system_and_oracle(ctx, small_batch):
    b = small_batch["batch"]
    for i, (s, tm) in enumerate(zip(small_batch["specs"], small_batch["tms"])):
        ph = sa.PhysicsParams.from_wavenumber(s["k"])
        beta = ph.burton_miller_beta()
        sys1 = sa.TbemAssemblyWorkspace(ctx, tm).assemble(ph, beta)
        A_single = sys1.matrix
        A_batch = b.matrix(i)
        assert np.array_equal(A_batch, A_single), f"instance {i}: batch matrix differs from single-system"
        # right-hand sides: the same device function, inlined with a constant unit amplitude
        B_single = sa.IncidentField(s["directions"]).compute_rhs_with_beta(tm, ph, beta).reshape(b.rhs(i).shape)
        assert rel_fro(b.rhs(i), B_single) < 1e-15
        c, nn, _ = tm.geometry()
        for q, d in enumerate(s["directions"]):
            ref = oracle.plane_wave_rhs(c, nn, s["k"], beta, direction=d)
            assert rel_l2(b.rhs(i)[:, q], ref) < 1e-14
        if i < 3:
            A_ref = oracle.build_tbem_system(s["mesh"].nodes, s["mesh"].triangles, s["k"], beta)[0]
            assert rel_fro(A_batch, A_ref) < 1e-12


def _check_solution(b, specs, sts, tol):
    col = 0
    for i, s in enumerate(specs):
        A = b.matrix(i)
        B = b.rhs(i)
        X = b.solution(i)
        for q in range(B.shape[1]):
            st = sts[col + q]
            res = np.linalg.norm(B[:, q] - A @ X[:, q]) / np.linalg.norm(B[:, q])
            assert st.converged, f"instance {i} column {q}: {st}"
            assert res < tol and abs(res - st.residual) <= 1e-10
            assert rel_l2(X[:, q], np.linalg.solve(A, B[:, q])) < 1e-7
        col += B.shape[1]


def test_batch_fp64_gmres(ctx, small_batch):
    b = small_batch["batch"]
    sts = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-8)
    _check_solution(b, small_batch["specs"], sts, 1e-8)
    # Arnoldi steps agree with the single-system GMRES(50) on the same system
    col = 0
    for i, s in enumerate(small_batch["specs"]):
        A_sys = sa.TbemAssemblyWorkspace(ctx, small_batch["tms"][i]).assemble(
            sa.PhysicsParams.from_wavenumber(s["k"]))
        op = sa.DenseMatrixOperator(A_sys)
        B = b.rhs(i)
        for q in range(B.shape[1]):
            single = sa.gmres(op, B[:, q], sa.GmresConfig(max_iterations=10, restart=50, tolerance=1e-8))
            assert abs(sts[col + q].iterations - single.iterations) <= 1, (i, q, sts[col + q], single)
        col += B.shape[1]


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_batch_refinement(ctx, small_batch, fmt):
    b = small_batch["batch"]
    b.convert(fmt)
    sts = b.solve(fmt + "_ir", restart=40, max_cycles=20, tolerance=1e-8, inner_tolerance=1e-4)
    _check_solution(b, small_batch["specs"], sts, 1e-8)
    print(fmt, "max refinement steps", max(s.cycles for s in sts), "max Arnoldi steps", max(s.iterations for s in sts))


def test_batch_graph_equals_direct(ctx, small_batch):
    b = small_batch["batch"]
    b.solve("fp64", restart=30, max_cycles=10, tolerance=1e-9, use_graph=False)
    direct = [b.solution(i).copy() for i in range(len(small_batch["specs"]))]
    b.solve("fp64", restart=30, max_cycles=10, tolerance=1e-9, use_graph=True)
    for i, x in enumerate(direct):
        assert np.array_equal(b.solution(i), x)
    # replaying the captured graph again gives the same result
    b.solve("fp64", restart=30, max_cycles=10, tolerance=1e-9, use_graph=True)
    for i, x in enumerate(direct):
        assert np.array_equal(b.solution(i), x)


def test_batch_autotune_keeps_results(ctx, small_batch):
    b = small_batch["batch"]
    before = [b.matrix(i).copy() for i in range(len(small_batch["specs"]))]
    best, table = b.autotune()
    assert len(table) == 16 and all(r["ms"] > 0 for r in table)
    assert best == b.tuning()
    b.assemble()
    for i, A in enumerate(before):
        assert np.array_equal(b.matrix(i), A)
    sts = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-8)
    _check_solution(b, small_batch["specs"], sts, 1e-8)


def test_batch_initial_guess(ctx, small_batch):
    """A guess near the solution converges in fewer Arnoldi steps to the same
    FP64 residual bar (measured relative to ||b||, as gmres_with_guess)."""
    b = small_batch["batch"]
    base = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-8)
    rng = np.random.default_rng(9)
    for i in range(len(small_batch["specs"])):
        X = b.solution(i)
        noise = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
        b.set_guess(i, X + 1e-4 * np.linalg.norm(X) / np.sqrt(X.size) * noise)
    sts = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-8, use_guess=True)
    _check_solution(b, small_batch["specs"], sts, 1e-8)
    assert sum(s.iterations for s in sts) < sum(s.iterations for s in base)
    # a zero guess reproduces the zero-start solve exactly
    for i in range(len(small_batch["specs"])):
        b.set_guess(i, np.zeros_like(b.solution(i)))
    again = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-8, use_guess=True)
    assert [s.iterations for s in again] == [s.iterations for s in base]


def test_batch_narrow_columns_fp64_path(ctx):
    """Batches whose instances all have <= 8 right-hand sides take the 8-column
    FP64 product; results match LAPACK and the single-system assembly."""
    rng = np.random.Generator(np.random.PCG64(21))
    specs = [{"mesh": ellipsoid(rng.uniform(0.6, 1.4, size=3), lvl), "k": k, "directions": uniform_directions(rng, nd)}
             for lvl, k, nd in ((1, 2.0, 1), (2, 4.0, 8), (1, 5.5, 3))]
    tms = [sa.TbemMesh(ctx, s["mesh"]) for s in specs]
    b = sa.Batch(ctx, tms)
    b.set_problem([s["k"] for s in specs], [s["directions"] for s in specs])
    b.assemble()
    sts = b.solve("fp64", restart=60, max_cycles=10, tolerance=1e-8)
    _check_solution(b, specs, sts, 1e-8)
    b.convert("fp16")
    sts = b.solve("fp16_ir", restart=40, max_cycles=20, tolerance=1e-8, inner_tolerance=1e-4)
    _check_solution(b, specs, sts, 1e-8)


def test_batch_apply_products(ctx, small_batch):
    """sbh_batch_apply: every instance's product with its own matrix in one
    launch.  FP64 agrees with the downloaded matrix times X to rounding; the
    FP16 and BF16 copies (per-row scaled) to their unit roundoff."""
    b = small_batch["batch"]
    rng = np.random.default_rng(11)
    X = [rng.standard_normal((s["mesh"].num_elements, int(c))) + 1j * rng.standard_normal((s["mesh"].num_elements, int(c)))
         for s, c in zip(small_batch["specs"], b.num_rhs)]
    ref = [b.matrix(i) @ x for i, x in enumerate(X)]
    Y = b.apply(X, "fp64")
    for y, r in zip(Y, ref):
        assert rel_fro(y, r) < 1e-14
    for fmt, tol in (("fp16", 2e-3), ("bf16", 1.5e-2)):
        b.convert(fmt)
        Y = b.apply(X, fmt)
        errs = [rel_fro(y, r) for y, r in zip(Y, ref)]
        print(fmt, "max relative error of the batched product", max(errs))
        assert max(errs) < tol
    with pytest.raises(sa.SbhError):
        b.apply(X, "fp16")  # the batch now holds the BF16 copy only
    Xn = [x.copy() for x in X]
    Xn[2][0, 0] = np.nan
    with pytest.raises(sa.SbhError):
        b.apply(Xn, "fp64")
