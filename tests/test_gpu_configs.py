"""GPU parity at the full sizes of the benchmark workloads.

The small-mesh tests pin every kernel against the oracle at up to 1280
panels.  Several size-dependent paths only show themselves at real size: the
uniform-grid near-pair search of the route builder (577,620 near pairs at
20480 panels), the split-K slabs of the products, the chunk counts of the
Krylov reductions and the batch work queues.  These tests run the HIP path
through the C ABI on each workload at its benchmark size:

* 20480-panel icosphere, k = 10, beta = i/k, +z, GMRES(100) — rows
  [0, 256) and 256 random rows of the FP64 matrix against the oracle's rows
  (build_tbem_system_with_beta_parallel, tbem.rs:1284-1466) to 1e-12 relative
  Frobenius and entrywise, with every near pair of those rows named by the
  oracle's own generate_subelements (singular.rs:1372-1528); the solve's
  recomputed FP64 residual, its agreement with the oracle's GMRES solution
  (committed fixture) and the surface pressure against the Mie series;
* 40960-panel Fibonacci sphere, k = 14, 64 Fibonacci plane waves — the same
  row sample, the FP64 multi-RHS solve with every column's residual, FP16 and
  BF16 refinement per column against the FP64 solution, and Mie per
  direction;
* 256 seeded ellipsoids — the whole batch, FP64 and FP16 refinement, every
  column's residual, 8 sampled instances against the oracle and LAPACK;
* the frequency sweep on the 10240-panel Fibonacci sphere — a reduced slice:
  solver-generated training data, training on the device, learned guesses in
  the batched FP16 refinement, and the zero-guess fallback.
"""

import os

import numpy as np
import pytest

import oracle
from oracle import mie
import sonium_amd as sa
from sonium_amd import warmstart as wsm
from sonium_amd.mesh import fibonacci_directions, fibonacci_sphere, icosphere, random_ellipsoid_batch

from helpers import rel_fro, rel_l2

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-12


def _sample_rows(n, seed, head=256, extra=256):
    rng = np.random.default_rng(seed)
    rest = rng.choice(np.arange(head, n), size=extra, replace=False)
    return np.concatenate([np.arange(head), np.sort(rest)])


def _gpu_rows(system, rows):
    """Rows of the device matrix in list order (contiguous runs in one copy)."""
    out = np.empty((rows.shape[0], system.num_dofs), dtype=np.complex128)
    i = 0
    while i < rows.shape[0]:
        j = i
        while j + 1 < rows.shape[0] and rows[j + 1] == rows[j] + 1:
            j += 1
        out[i: j + 1] = system.matrix_rows(int(rows[i]), int(rows[j]) + 1)
        i = j + 1
    return out


def _near_columns(m, rows):
    """Near pairs (i, j != i) of the sample as the oracle's generate_subelements
    decides them: a pair is near when the subdivision returns more than the one
    full leaf.  Only columns within 4 sqrt(area_j) of the collocation point can
    be near (the leaf test at level 0 is dist / sqrt(area) >= 3)."""
    c, _, area = m.element_geometry()
    tri = m.triangles
    out = []
    for i in rows:
        d = np.linalg.norm(c - c[i], axis=1)
        cand = np.nonzero(d < 4.0 * np.sqrt(area))[0]
        for j in cand:
            if j == i:
                continue
            coords = m.nodes[tri[j]].reshape(9)
            if oracle.generate_subelements(c[i], coords, area[j]).shape[0] > 1:
                out.append((int(i), int(j)))
    return out


def _check_rows(m, system, k, beta, rows, label):
    A_gpu = _gpu_rows(system, rows)
    A_ref, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, beta, rows=rows)
    err = rel_fro(A_gpu, A_ref)
    diff = np.abs(A_gpu - A_ref)
    scale = np.abs(A_ref).max(axis=1, keepdims=True)
    entry = float(np.max(diff / (np.abs(A_ref) + 1e-6 * scale)))
    near = _near_columns(m, rows)
    pos = {int(r): q for q, r in enumerate(rows)}
    near_err = max(float(abs(A_gpu[pos[i], j] - A_ref[pos[i], j]) / abs(A_ref[pos[i], j])) for i, j in near)
    print(f"{label}: {rows.shape[0]} rows, rel Frobenius {err:.3e}, worst entry {entry:.3e}, "
          f"{len(near)} near pairs (worst {near_err:.3e})")
    assert err <= TOL
    # a single flipped near/far decision changes its entry by ~1e-4 relative
    assert entry <= 1e-9
    assert len(near) >= 10 * rows.shape[0] and near_err <= 1e-11
    return A_ref


# --------------------------------------------------------------------------- 20480 panels


@pytest.fixture(scope="module")
def sphere20480(ctx):
    m = icosphere(1.0, 5)
    k = 10.0
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    system = ws.assemble(ph, beta)
    sa.IncidentField.plane_wave_z().add_to_system(system, tm, ph, beta, sign=1.0)
    yield {"mesh": m, "tm": tm, "ph": ph, "beta": beta, "ws": ws, "sys": system, "k": k}
    system.close()
    ws.close()


def test_c2_row_sample_matches_oracle(sphere20480):
    s = sphere20480
    m = s["mesh"]
    st = s["ws"].stats
    assert st["near_pairs"] == 577620  # the route builder's count at this mesh
    rows = _sample_rows(m.num_elements, seed=20480)
    _check_rows(m, s["sys"], s["k"], s["beta"], rows, "20480 panels")
    c, nn, _ = m.element_geometry()
    b_ref = oracle.plane_wave_rhs(c, nn, s["k"], s["beta"])
    assert rel_l2(s["sys"].rhs, b_ref) <= 1e-14


def test_c2_solve_matches_oracle_solution_and_mie(sphere20480):
    s = sphere20480
    m = s["mesh"]
    op = sa.DenseMatrixOperator(s["sys"])
    b = s["sys"].rhs
    sol = sa.gmres(op, b, sa.GmresConfig(max_iterations=100, restart=100, tolerance=1e-8))
    res = float(np.linalg.norm(b - op.apply(sol.x)) / np.linalg.norm(b))
    fx = np.load(os.path.join(GOLDEN, "oracle_solution_icosphere20480_k10.npz"))
    err_x = rel_l2(sol.x, fx["x"])
    c, _, area = m.element_geometry()
    err_mie = mie.weighted_relative_l2(sol.x, mie.surface_pressure(s["k"], 1.0, c), area)
    print(f"20480 panels: {sol.iterations} steps (oracle {int(fx['iterations'])}), residual {sol.residual:.3e} "
          f"(recomputed {res:.3e}), vs oracle solution {err_x:.3e}, surface L2 vs Mie {err_mie:.4e} "
          f"(oracle {float(fx['mie_l2']):.4e})")
    assert sol.converged and sol.residual <= 1e-8 and res <= 1e-8
    assert abs(sol.iterations - int(fx["iterations"])) <= 1
    assert err_x <= 1e-8
    # same discretisation as the oracle: the same distance from the series
    assert abs(err_mie - float(fx["mie_l2"])) <= 1e-6 and err_mie < 0.02
    # residual of the GPU solution against the oracle's own rows (independent of the GPU product)
    rows = _sample_rows(m.num_elements, seed=7, head=0, extra=256)
    A_ref, _ = oracle.build_tbem_system(m.nodes, m.triangles, s["k"], s["beta"], rows=rows)
    r = b[rows] - A_ref @ sol.x
    assert np.linalg.norm(r) / np.linalg.norm(b[rows]) <= 1e-7


# --------------------------------------------------------------------------- 40960 panels, 64 RHS


@pytest.fixture(scope="module")
def hull40960(ctx):
    m = fibonacci_sphere(1.0, 129, 160)
    assert m.num_elements == 40960
    k = 14.0
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    system = ws.assemble(ph, beta)
    dirs = fibonacci_directions(64)
    n = m.num_elements
    B = sa.DeviceVector(ctx, n * 64)
    sa.IncidentField(dirs).write_device(B, tm, ph, beta)
    yield {"mesh": m, "tm": tm, "ph": ph, "beta": beta, "ws": ws, "sys": system, "k": k, "dirs": dirs, "B": B}
    B.close()
    system.close()
    ws.close()


def _block_residuals(op, B, X):
    R = B - op.apply_block(X)
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


def test_c3_row_sample_matches_oracle(hull40960):
    s = hull40960
    rows = _sample_rows(s["mesh"].num_elements, seed=40960)
    _check_rows(s["mesh"], s["sys"], s["k"], s["beta"], rows, "40960 panels")


def test_c3_fp64_and_low_precision_solves(hull40960, ctx):
    s = hull40960
    m = s["mesh"]
    n = m.num_elements
    op = sa.DenseMatrixOperator(s["sys"])
    X = sa.DeviceVector(ctx, n * 64)
    sols = sa.gmres_block_device(op, s["B"], 64, X, sa.GmresConfig(max_iterations=100, restart=100, tolerance=1e-8))
    Bh = s["B"].download().reshape(n, 64)
    X64 = X.download().reshape(n, 64)
    res = _block_residuals(op, Bh, X64)
    print(f"40960 panels FP64: max steps {max(q.iterations for q in sols)}, max residual {res.max():.3e}")
    assert all(q.converged and q.residual <= 1e-8 for q in sols)
    assert res.max() <= 1e-8
    c, _, area = m.element_geometry()
    for col in (0, 21, 42, 63):
        p = mie.surface_pressure(s["k"], 1.0, c, direction=s["dirs"][col])
        err = mie.weighted_relative_l2(X64[:, col], p, area)
        print(f"  direction {col}: surface L2 vs Mie {err:.4e}")
        assert err < 0.03
    cfg = sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8, max_inner_iterations=50)
    for fmt, itol in (("fp16", 1e-3), ("bf16", 1e-2)):
        lp = sa.LowpMatrix(s["sys"], fmt)
        cfg.inner_tolerance = itol
        st = sa.gmres_ir_device(op, lp, s["B"], 64, X, cfg)
        Xl = X.download().reshape(n, 64)
        res_l = _block_residuals(op, Bh, Xl)
        err = np.linalg.norm(Xl - X64, axis=0) / np.linalg.norm(X64, axis=0)
        print(f"40960 panels {fmt}: max outer {max(q.outer_iterations for q in st)}, max residual "
              f"{res_l.max():.3e}, max error vs FP64 {err.max():.3e}")
        assert all(q.converged and q.final_outer_residual <= 1e-8 for q in st)
        assert res_l.max() <= 1e-8
        # both solutions meet 1e-8; their distance is bounded by the conditioning
        assert err.max() <= 1e-6
        lp.close()
    X.close()


def test_c3_shared_basis_block_gmres(hull40960, ctx):
    """The reference's multi-RHS solver (block_gmres_inner, block_gmres.rs:666-1120)
    at full size: every column's recomputed FP64 residual, Mie per direction,
    and the per-column GMRES solution within the conditioning bound."""
    import time

    s = hull40960
    m = s["mesh"]
    n = m.num_elements
    op = sa.DenseMatrixOperator(s["sys"])
    X = sa.DeviceVector(ctx, n * 64)
    cfg = sa.BlockGmresConfig(max_iterations=100, restart=100, tolerance=1e-8)
    ctx.synchronize()
    t0 = time.perf_counter()
    sol = sa.block_gmres_device(op, s["B"], 64, X, cfg)
    el = time.perf_counter() - t0
    Bh = s["B"].download().reshape(n, 64)
    Xb = X.download().reshape(n, 64)
    res = _block_residuals(op, Bh, Xb)
    print(f"40960 panels block GMRES: {sol.iterations} block steps, basis {sol.max_basis}, {sol.deflated} deflated, "
          f"{sol.spmv_count} column products, {el:.3f} s, max residual {res.max():.3e} "
          f"(reported {sol.true_residuals.max():.3e})")
    assert sol.converged and sol.true_residuals.max() <= 1e-8 and res.max() <= 1e-8
    np.testing.assert_allclose(sol.true_residuals, res, rtol=1e-6, atol=1e-14)
    c, _, area = m.element_geometry()
    for col in (0, 31, 63):
        p = mie.surface_pressure(s["k"], 1.0, c, direction=s["dirs"][col])
        err = mie.weighted_relative_l2(Xb[:, col], p, area)
        assert err < 0.03
    X2 = sa.DeviceVector(ctx, n * 64)
    sa.gmres_block_device(op, s["B"], 64, X2, sa.GmresConfig(max_iterations=100, restart=100, tolerance=1e-8))
    X64 = X2.download().reshape(n, 64)
    err = np.linalg.norm(Xb - X64, axis=0) / np.linalg.norm(X64, axis=0)
    print(f"  vs per-column GMRES: max {err.max():.3e}")
    assert err.max() <= 1e-6
    X.close()
    X2.close()


# --------------------------------------------------------------------------- 256 ellipsoids


def test_c4_full_batch(ctx):
    specs = random_ellipsoid_batch(256, seed=20261202)
    tms = [sa.TbemMesh(ctx, sp["mesh"]) for sp in specs]
    batch = sa.Batch(ctx, tms)
    batch.set_problem([sp["k"] for sp in specs], [sp["directions"] for sp in specs])
    batch.assemble()
    sts64 = batch.solve("fp64", restart=100, max_cycles=10, tolerance=1e-8, use_graph=True)
    ncol = np.array([len(sp["directions"]) for sp in specs])
    assert len(sts64) == int(ncol.sum())
    assert all(q.converged and q.residual <= 1e-8 for q in sts64)
    X64 = {i: batch.solution(i) for i in range(0, 256, 32)}
    batch.convert("fp16")
    sts16 = batch.solve("fp16_ir", restart=30, max_cycles=30, tolerance=1e-8, inner_tolerance=1e-3, use_graph=True)
    assert all(q.converged and q.residual <= 1e-8 for q in sts16)
    print(f"256 instances, {len(sts64)} columns: FP64 max residual {max(q.residual for q in sts64):.3e}, "
          f"FP16 refinement max residual {max(q.residual for q in sts16):.3e}")
    worst = 0.0
    for i in range(0, 256, 32):
        sp = specs[i]
        msh = sp["mesh"]
        beta = 1j / sp["k"]
        A_ref, _ = oracle.build_tbem_system(msh.nodes, msh.triangles, sp["k"], beta)
        A = batch.matrix(i)
        assert rel_fro(A, A_ref) <= TOL
        c, nn, _ = msh.element_geometry()
        B_ref = np.stack([oracle.plane_wave_rhs(c, nn, sp["k"], beta, direction=d) for d in sp["directions"]], 1)
        assert rel_fro(batch.rhs(i), B_ref) <= 1e-14
        X_lu = np.linalg.solve(A_ref, B_ref)
        X16 = batch.solution(i)
        for q in range(B_ref.shape[1]):
            r64 = np.linalg.norm(B_ref[:, q] - A_ref @ X64[i][:, q]) / np.linalg.norm(B_ref[:, q])
            r16 = np.linalg.norm(B_ref[:, q] - A_ref @ X16[:, q]) / np.linalg.norm(B_ref[:, q])
            assert r64 <= 1e-8 and r16 <= 1e-8
            assert rel_l2(X64[i][:, q], X_lu[:, q]) <= 1e-7
            assert rel_l2(X16[:, q], X_lu[:, q]) <= 1e-7
            worst = max(worst, r64, r16)
        print(f"  instance {i}: {msh.num_elements} panels, k={sp['k']:.3f}, {B_ref.shape[1]} directions")
    print(f"sampled instances: worst residual against the oracle matrix {worst:.3e}")
    batch.close()
    for tm in tms:
        tm.close()


# --------------------------------------------------------------------------- sweep slice


def test_c5_sweep_slice(ctx):
    """A reduced slice of the sweep: 24 training wavenumbers in [1, 12] on the
    10240-panel sphere, 400 Adam steps on the device, guesses for 8 held-out
    wavenumbers solved by the batched FP16 refinement to the FP64 residual,
    and the zero-guess fallback for a guess that is worse than nothing."""
    m = fibonacci_sphere(1.0, 65, 80)
    n = m.num_elements
    assert n == 10240
    tm = sa.TbemMesh(ctx, m)
    c, nn, _ = tm.geometry()
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    cfg = sa.GmresConfig(max_iterations=20, restart=100, tolerance=1e-8)
    d = np.array([0.0, 0.0, 1.0])
    inputs, targets = [], []
    for k in np.linspace(1.0, 12.0, 24):
        ph = sa.PhysicsParams.from_wavenumber(float(k))
        beta = ph.burton_miller_beta()
        ws.assemble(ph, beta, system)
        b = sa.IncidentField([d]).compute_rhs_with_beta(tm, ph, beta)
        sol = sa.gmres(op, b, cfg)
        assert sol.converged
        inputs.append(wsm.panel_inputs(c, nn, float(k), d))
        targets.append(wsm.density_to_target(sol.x, c, float(k), d))
    net = wsm.WarmStartMLP(ctx, width=512, seed=11)
    net.set_dataset(np.concatenate(inputs), np.concatenate(targets))
    losses = net.train(400, batch=16384, lr_max=2e-3, lr_min=2e-5, seed=3)
    assert np.all(np.isfinite(losses)) and np.mean(losses[-20:]) < 0.5 * np.mean(losses[:20])

    kh = np.linspace(1.0, 12.0, 24)[:8] + 0.5 * (12.0 - 1.0) / 23
    X0 = net.guesses(c, nn, kh, d)
    batch = sa.Batch(ctx, [tm] * len(kh))
    batch.set_problem([float(k) for k in kh], [d.reshape(1, 3)] * len(kh))
    batch.assemble()
    batch.convert("fp16")
    for i in range(len(kh)):
        batch.set_guess(i, X0[i].reshape(n, 1))
    sts = batch.solve("fp16_ir", restart=60, max_cycles=20, tolerance=1e-8, inner_tolerance=1e-4, use_guess=True)
    assert all(q.converged and q.residual <= 1e-8 for q in sts)
    for i, k in enumerate(kh):
        ph = sa.PhysicsParams.from_wavenumber(float(k))
        beta = ph.burton_miller_beta()
        ws.assemble(ph, beta, system)
        b = sa.IncidentField([d]).compute_rhs_with_beta(tm, ph, beta)
        x = batch.solution(i)[:, 0]
        assert np.linalg.norm(b - op.apply(x)) / np.linalg.norm(b) <= 1e-8
        ref = sa.gmres(op, b, cfg)
        assert rel_l2(x, ref.x) <= 1e-6
        if i == 0:
            # the fallback: a guess whose residual exceeds ||b|| is dropped for the zero guess
            bad = 1e3 * (np.random.default_rng(1).standard_normal(n) + 0j)
            sol, used, r0 = wsm.solve_with_guess(op, b, bad, cfg, with_residual=True)
            assert not used and r0 > 1.0 and sol.converged and sol.iterations == ref.iterations
            sol, used, r0 = wsm.solve_with_guess(op, b, X0[i], cfg, with_residual=True)
            print(f"k={k:.3f}: learned guess residual {r0:.3e}, {sol.iterations} steps (zero guess "
                  f"{ref.iterations})")
            assert used and r0 < 1.0 and sol.converged
    batch.close()
    net.close()
    system.close()
    ws.close()
    tm.close()
