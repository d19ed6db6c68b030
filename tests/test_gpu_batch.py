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
def test_batch_lowp_copy_is_bit_exact(ctx, small_batch, fmt):
    """The batch conversion (one row per workgroup) equals the oracle's
    row-scaled storage bit for bit, for every instance size."""
    from oracle import lowp as olp

    b = small_batch["batch"]
    b.convert(fmt)
    for i in range(len(small_batch["specs"])):
        halves, inv, ld = b.lowp_copy(i)
        ref_h, ref_inv = olp.row_scaled_halves(b.matrix(i), fmt, ld)
        assert np.array_equal(inv, ref_inv), f"instance {i}: row scales differ"
        assert np.array_equal(halves, ref_h), f"instance {i}: {np.count_nonzero(halves != ref_h)} halves differ"


def test_batch_lowp_copy_long_rows(ctx):
    """Rows longer than a conversion block holds in registers (12800 panels)
    take the two-pass conversion; sampled rows equal the oracle bit for bit."""
    from oracle import lowp as olp
    from sonium_amd.mesh import fibonacci_sphere

    tm = sa.TbemMesh(ctx, fibonacci_sphere(1.0, 65, 100))
    b = sa.Batch(ctx, [tm])
    try:
        b.set_problem([3.0], [np.array([[0.0, 0.0, 1.0]])])
        b.assemble()
        b.convert("fp16")
        halves, inv, ld = b.lowp_copy(0)
        A = b.matrix(0)
        rows = np.unique(np.concatenate([[0, len(A) - 1], np.random.Generator(np.random.PCG64(3)).integers(
            0, len(A), 256)]))
        ref_h, ref_inv = olp.row_scaled_halves(A[rows], "fp16", ld)
        assert np.array_equal(inv[rows], ref_inv)
        assert np.array_equal(halves[rows], ref_h)
    finally:
        b.close()
        tm.close()


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
    # packed transfers equal the per-instance ones
    for i, r in enumerate(b.rhs_all()):
        assert np.array_equal(r, b.rhs(i))
    b.set_guesses(X)
    for i, x in enumerate(b.solutions()):
        assert np.array_equal(x, X[i]) and np.array_equal(b.solution(i), X[i])
    with pytest.raises(sa.SbhError):
        b.set_guesses(Xn)


def test_batch_sweep_guesses_from_solved_neighbours(ctx):
    """Interleaved sweep groups: the second group starts from minimal-residual
    combinations of the first group's solutions (warmstart.batch_guesses),
    whose residuals, formed from batched products on the device, match the
    host recomputation, and which converge in fewer Arnoldi steps than zero
    starts; "previous" takes the nearest solved point below."""
    from sonium_amd import warmstart as wsm
    from sonium_amd.mesh import icosphere

    m = icosphere(1.0, 2)
    tm = sa.TbemMesh(ctx, m)
    c, nn, _ = tm.geometry()
    ks = np.linspace(2.0, 2.7, 8)
    d = np.array([0.0, 0.0, 1.0])
    b = sa.Batch(ctx, [tm] * 4)
    solved = {}
    idx = [0, 2, 4, 6]
    b.set_problem([float(ks[i]) for i in idx], [d.reshape(1, 3)] * 4)
    b.assemble()
    st = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-10, use_graph=False)
    assert all(s.converged for s in st)
    for i, j in enumerate(idx):
        solved[j] = b.solution(i)[:, 0].copy()
    idx = [1, 3, 5, 7]
    b.set_problem([float(ks[i]) for i in idx], [d.reshape(1, 3)] * 4)
    b.assemble()
    zero = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-10, use_graph=False)
    X0, rels = wsm.batch_guesses(b, "neighbours", idx, solved, None, c, nn, ks[idx], d)
    for i in range(4):
        A, rhs = b.matrix(i), b.rhs(i)[:, 0]
        r = np.linalg.norm(rhs - A @ X0[i]) / np.linalg.norm(rhs)
        assert abs(r - rels[i]) <= 1e-12 + 1e-8 * r
        assert r < (0.05 if i < 3 else 0.2)  # the last point has a solved neighbour below only
    guessed = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-10, use_graph=False, use_guess=True)
    print("zero-start steps", [s.iterations for s in zero], "neighbour-combination steps",
          [s.iterations for s in guessed], "guess residuals", ["%.1e" % r for r in rels])
    assert all(s.converged for s in guessed)
    assert sum(s.iterations for s in guessed) < sum(s.iterations for s in zero)
    for i in range(4):
        A, rhs = b.matrix(i), b.rhs(i)[:, 0]
        assert np.linalg.norm(rhs - A @ b.solution(i)[:, 0]) / np.linalg.norm(rhs) < 1e-10
    Xp, _ = wsm.batch_guesses(b, "previous", idx, solved, None, c, nn, ks[idx], d)
    for i, j in enumerate(idx):
        assert np.array_equal(Xp[i], solved[j - 1])
    b.close()
    tm.close()


def test_device_guesses_match_host_guesses(ctx):
    """The device-side guess pipeline of the sweep driver gives the host
    pipeline's guesses: solutions stored on the device and set back as guesses
    bit for bit; previous-point guesses bit for bit; minimal-residual
    combinations from sbh_batch_guess_gram / sbh_batch_guess_combine equal
    warmstart.batch_guesses (host products and normal equations) to 1e-12,
    and their residuals, formed from the Gram data, equal the host
    recomputation b - A x0; the network's guesses (sbh_batch_learned_guesses)
    equal its predictions on the same inputs times exp(i k d.c) to 1e-14.
    Host addresses in a pointer table are refused before any launch."""
    from sonium_amd import warmstart as wsm
    from sonium_amd.mesh import icosphere

    m = icosphere(1.0, 2)
    n = m.num_elements
    tm = sa.TbemMesh(ctx, m)
    c, nn, _ = tm.geometry()
    ks = np.linspace(2.0, 2.7, 8)
    d = np.array([0.0, 0.0, 1.0])
    b = sa.Batch(ctx, [tm] * 4)
    store = sa.DeviceVector(ctx, len(ks) * n)
    ptr = lambda j: store.ptr + 16 * n * j  # noqa: E731
    idx = [0, 2, 4, 6]
    b.set_problem([float(ks[i]) for i in idx], [d.reshape(1, 3)] * 4)
    b.assemble()
    st = b.solve("fp64", restart=50, max_cycles=10, tolerance=1e-10, use_graph=False)
    assert all(s.converged for s in st)
    solved = {j: b.solution(i)[:, 0].copy() for i, j in enumerate(idx)}
    b.store_solutions([ptr(j) for j in idx])
    for j in idx:
        assert np.array_equal(store.download(j * n, n), solved[j])
    b.set_guesses_device([ptr(j) for j in idx[::-1]])
    assert all(np.array_equal(x[:, 0], solved[j]) for x, j in zip(b.solutions(), idx[::-1]))

    idx = [1, 3, 5, 7]
    b.set_problem([float(ks[i]) for i in idx], [d.reshape(1, 3)] * 4)
    b.assemble()
    b.convert("fp16")
    wsm.batch_guesses_device(b, "previous", idx, set(solved), ptr)
    for x, j in zip(b.solutions(), idx):
        assert np.array_equal(x[:, 0], solved[j - 1])
    for fmt in ("fp64", "fp16"):
        X0, rels = wsm.batch_guesses(b, "neighbours", idx, solved, None, c, nn, ks[idx], d, products=fmt)
        rels_dev = wsm.batch_guesses_device(b, "neighbours", idx, set(solved), ptr, products=fmt)
        Xd = [x[:, 0] for x in b.solutions()]
        for i in range(4):
            assert rel_l2(Xd[i], X0[i]) < 1e-12
            A, rhs = b.matrix(i), b.rhs(i)[:, 0]
            r = np.linalg.norm(rhs - A @ Xd[i]) / np.linalg.norm(rhs)
            if fmt == "fp64":
                assert abs(rels_dev[i] - r) <= 1e-12 + 1e-8 * r
            assert abs(rels_dev[i] - rels[i]) <= 1e-12 + 1e-8 * rels[i]
    # a host address is refused with an error, not passed to a kernel
    host = np.zeros(n, dtype=np.complex128)
    with pytest.raises(sa.SbhError):
        b.set_guesses_device([host.ctypes.data, None, None, None])

    # a new problem invalidates the gathered candidates
    b.set_problem([float(ks[i]) for i in idx], [d.reshape(1, 3)] * 4)
    with pytest.raises(sa.SbhError):
        b.guess_combine(np.ones((4, 2), dtype=np.complex128))
    b.assemble()

    net = wsm.WarmStartMLP(ctx, width=512, seed=5)
    b.learned_guesses(net)
    Xl = np.stack([x[:, 0] for x in b.solutions()])
    raw = np.concatenate([wsm.panel_inputs(c, nn, float(ks[j]), d) for j in idx])
    y = net.predict(raw).astype(np.float64).reshape(4, n, 2)
    ref = (y[..., 0] + 1j * y[..., 1]) * np.exp(1j * ks[idx][:, None] * (c @ d)[None, :])
    assert np.max(np.abs(Xl - ref)) <= 1e-14 * np.max(np.abs(ref))
    buf = sa.DeviceVector(ctx, 4 * n)
    b.learned_guesses(net, out=buf)
    assert np.array_equal(buf.download().reshape(4, n), Xl)
    rels_l = wsm.batch_guesses_device(b, "neighbours+learned", idx, set(solved), ptr, net=net, learned_buffer=buf)
    assert all(0.0 < r < 1.0 for r in rels_l)
    net.close()
    buf.close()
    store.close()
    b.close()
    tm.close()


def test_frequency_family_matches_single_system(ctx):
    """Instances on one mesh with equally spaced wavenumbers share one pass over
    the pair geometry (batch_far_family_kernel for the far entries,
    batch_near_family_kernel for the subdivided near pairs; phases advanced by
    complex products).  Their matrices agree with the single-system assembly
    to rounding (1e-13 relative Frobenius and 1e-13 of the largest entry per
    entry, the recurrence and a different summation order) and with the
    oracle to the 1e-12 bar; the
    right-hand sides are unchanged; an instance without an equally spaced
    partner keeps the per-instance loop, bit for bit."""
    from sonium_amd.mesh import icosphere

    m = icosphere(1.0, 3)
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    d = np.array([[0.0, 0.0, 1.0]])
    ks = 2.0 + 0.05 * np.arange(20)
    perm = np.random.default_rng(4).permutation(20)  # the batch order need not be sorted
    b = sa.Batch(ctx, [tm] * 20)
    b.set_problem([float(ks[p]) for p in perm], [d] * 20)
    b.assemble()
    worst = 0.0
    for i in range(20):
        k = float(ks[perm[i]])
        ph = sa.PhysicsParams.from_wavenumber(k)
        beta = ph.burton_miller_beta()
        A1 = ws.assemble(ph, beta).matrix
        Ab = b.matrix(i)
        err = rel_fro(Ab, A1)
        worst = max(worst, err)
        assert err < 1e-13, f"instance {i} (k={k}): {err:.2e}"
        ent = np.max(np.abs(Ab - A1)) / np.max(np.abs(A1))  # the near pairs hold the largest entries
        assert ent < 1e-13, f"instance {i} (k={k}): entrywise {ent:.2e}"
        if i in (0, 19):
            B1 = sa.IncidentField(d).compute_rhs_with_beta(tm, ph, beta).reshape(-1, 1)
            assert rel_fro(b.rhs(i), B1) < 1e-15
            A_ref = oracle.build_tbem_system(m.nodes, m.triangles, k, beta)[0]
            assert rel_fro(b.matrix(i), A_ref) < 1e-12
    print(f"frequency families (16 + 4 members): worst relative Frobenius difference {worst:.2e}")
    sts = b.solve("fp64", restart=60, max_cycles=5, tolerance=1e-8)
    assert all(s.converged and s.residual < 1e-8 for s in sts)
    # irregular spacing: 2.0 and 2.13 form a family of two, 2.5 is left alone and keeps
    # the per-instance loop, bit for bit
    b2 = sa.Batch(ctx, [tm] * 3)
    kk = [2.0, 2.13, 2.5]
    b2.set_problem(kk, [d] * 3)
    b2.assemble()
    for i, k in enumerate(kk):
        ph = sa.PhysicsParams.from_wavenumber(k)
        A1 = ws.assemble(ph, ph.burton_miller_beta()).matrix
        if i == 2:
            assert np.array_equal(b2.matrix(i), A1)
        else:
            assert rel_fro(b2.matrix(i), A1) < 1e-13
    b.close()
    b2.close()
    ws.close()
    tm.close()
