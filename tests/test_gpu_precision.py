"""GPU tests of the many-right-hand-side and mixed-precision paths.

* the complex FP16/BF16 copy: per-row scales are exact powers of two that
  bracket the row maximum, and every entry is the FP64 entry rounded once to
  the format (relative error <= 2^-11 for FP16, 2^-8 for BF16, plus the
  subnormal floor of the scaled row);
* low-precision products (GEMV for one column, MFMA GEMM for many) against
  numpy on the dequantized operands (FP32 accumulation), and against the FP64
  product within the format's rounding;
* block GMRES: every column equals the single right-hand-side GMRES
  (gmres.rs:631-1057) — same iteration count, solution to 1e-10;
* FGMRES-IR (gmres_ir.rs:665): recomputed FP64 residual <= 1e-8 and the
  solution within 1e-7 of the FP64 oracle (LAPACK on the oracle matrix),
  reported per right-hand side.
"""

import numpy as np
import pytest

import oracle
import sonium_amd as sa
from sonium_amd.mesh import fibonacci_directions, icosphere

from helpers import rel_fro, rel_l2

pytestmark = pytest.mark.gpu

U = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8}


@pytest.fixture(scope="module")
def c1_block(ctx):
    m = icosphere(1.0, 3)
    k = 2.0
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    sys = ws.assemble(ph, beta)
    dirs = fibonacci_directions(12)
    B = sa.IncidentField(dirs).compute_rhs_with_beta(tm, ph, beta)
    return {"mesh": m, "tm": tm, "sys": sys, "A": sys.matrix, "B": B, "ws": ws, "beta": beta}


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_lowp_copy_rounding_and_power_of_two_scales(c1_block, fmt):
    A = c1_block["A"]
    lp = sa.LowpMatrix(c1_block["sys"], fmt)
    halves, inv, ld = lp.download()
    assert ld % 64 == 0 and ld >= A.shape[0]
    top = 14 if fmt == "fp16" else 0
    rowmax = np.maximum(np.abs(A.real), np.abs(A.imag)).max(axis=1)
    e = np.log2(inv) + top
    assert np.all(e == np.round(e))  # exact powers of two
    assert np.all(2.0 ** e <= rowmax) and np.all(rowmax < 2.0 ** (e + 1))
    D = lp.dequantized()
    floor = (2.0 ** -24 if fmt == "fp16" else 0.0) * inv[:, None]  # FP16 subnormal spacing, unscaled
    err_re = np.abs(D.real - A.real)
    err_im = np.abs(D.imag - A.imag)
    # one rounding to the format (plus the FP64->FP32->format double rounding, < 2^-24)
    assert np.all(err_re <= 1.01 * U[fmt] * np.abs(A.real) + floor + 1e-300)
    assert np.all(err_im <= 1.01 * U[fmt] * np.abs(A.imag) + floor + 1e-300)
    # padding columns are zero
    assert np.all(halves[:, 2 * A.shape[0]:] == 0)


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
@pytest.mark.parametrize("k", [1, 5, 17, 64])
def test_lowp_products(c1_block, fmt, k):
    A = c1_block["A"]
    n = A.shape[0]
    lp = sa.LowpMatrix(c1_block["sys"], fmt)
    D = lp.dequantized()
    rng = np.random.default_rng(k)
    X = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    Y = lp.apply(X if k > 1 else X[:, 0]).reshape(n, k)
    # the operand is quantized per column with a power-of-two scale; the product
    # is then the dequantized one up to FP32 accumulation
    from oracle import lowp as olp

    Xq = olp.quantized_columns(X, fmt)
    assert rel_fro(Y, D @ Xq) < 5e-6
    assert rel_fro(Y, D @ X) < 4 * U[fmt]
    assert rel_fro(Y, A @ X) < 6 * U[fmt]
    # each column equals the single-column product of its own quantized vector
    col = 3 % k
    ycol = lp.apply(X[:, col].copy())
    assert rel_l2(Y[:, col], ycol) < 5e-6


def test_fp64_block_apply(c1_block):
    op = sa.DenseMatrixOperator(c1_block["sys"])
    A = c1_block["A"]
    rng = np.random.default_rng(4)
    for k in (2, 7, 64, 70):
        X = rng.standard_normal((A.shape[0], k)) + 1j * rng.standard_normal((A.shape[0], k))
        assert rel_fro(op.apply_block(X), A @ X) < 1e-14


def test_block_gmres_columns_equal_single_gmres(c1_block):
    op = sa.DenseMatrixOperator(c1_block["sys"])
    B = c1_block["B"]
    cfg = sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    X, sols = sa.gmres_block(op, B, cfg)
    for c in range(B.shape[1]):
        single = sa.gmres(op, B[:, c], cfg)
        assert sols[c].converged and sols[c].residual <= 1e-8
        assert sols[c].iterations == single.iterations
        assert rel_l2(X[:, c], single.x) < 1e-10
        xr, st = oracle.gmres(c1_block["A"], B[:, c], restart=50, max_iterations=100, tol=1e-8)
        assert st["iterations"] == sols[c].iterations
        assert rel_l2(X[:, c], xr) < 1e-8


@pytest.mark.parametrize("k", [1, 64])
def test_block_gmres_one_and_64_columns(c1_block, k):
    """One column runs the short-chunk Krylov kernels (at most 2 rows per thread,
    the projection unrolled over 4 basis vectors); 64 columns run the wide ones
    (8 rows per thread).  Each column follows the single-RHS GMRES either way."""
    op = sa.DenseMatrixOperator(c1_block["sys"])
    ph = sa.PhysicsParams.from_wavenumber(2.0)
    B = sa.IncidentField(fibonacci_directions(64)).compute_rhs_with_beta(c1_block["tm"], ph, c1_block["beta"])
    B = B[:, :k].copy()
    cfg = sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    X, sols = sa.gmres_block(op, B, cfg)
    for c in sorted({0, k // 2, k - 1}):
        single = sa.gmres(op, B[:, c], cfg)
        assert sols[c].converged and sols[c].residual <= 1e-8
        assert abs(sols[c].iterations - single.iterations) <= 1
        assert rel_l2(X[:, c], single.x) < 1e-9
    R = B - c1_block["A"] @ X
    assert np.all(np.linalg.norm(R, axis=0) <= 1e-8 * np.linalg.norm(B, axis=0) * 1.01)


def test_block_gmres_short_restart_matches_single(c1_block):
    op = sa.DenseMatrixOperator(c1_block["sys"])
    B = c1_block["B"][:, :5]
    cfg = sa.GmresConfig(max_iterations=60, restart=6, tolerance=1e-9)
    X, sols = sa.gmres_block(op, B, cfg)
    for c in range(B.shape[1]):
        single = sa.gmres(op, B[:, c], cfg)
        assert sols[c].converged == single.converged
        assert sols[c].restarts == single.restarts
        assert abs(sols[c].iterations - single.iterations) <= 1
        assert rel_l2(X[:, c], single.x) < 1e-7


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_gmres_ir_reaches_fp64_accuracy(c1_block, fmt):
    op = sa.DenseMatrixOperator(c1_block["sys"])
    lp = sa.LowpMatrix(c1_block["sys"], fmt)
    B = c1_block["B"]
    cfg = sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8, max_inner_iterations=50,
                           inner_tolerance=1e-3 if fmt == "fp16" else 1e-2)
    X, st = sa.gmres_ir(op, lp, B, cfg)
    A = c1_block["A"]
    for c in range(B.shape[1]):
        xl = np.linalg.solve(A, B[:, c])
        err = rel_l2(X[:, c], xl)
        print(f"{fmt} rhs {c}: outer {st[c].outer_iterations} inner {st[c].total_inner_iterations} "
              f"residual {st[c].final_outer_residual:.2e} error vs FP64 {err:.2e}")
        assert st[c].converged and st[c].final_outer_residual <= 1e-8
        assert np.linalg.norm(B[:, c] - A @ X[:, c]) / np.linalg.norm(B[:, c]) <= 1e-8
        assert err < 1e-7
        assert st[c].outer_iterations <= 10


def test_gmres_ir_single_rhs_gemv_path(c1_block):
    op = sa.DenseMatrixOperator(c1_block["sys"])
    lp = sa.LowpMatrix(c1_block["sys"], "fp16")
    b = c1_block["B"][:, 0]
    X, st = sa.gmres_ir(op, lp, b, sa.GmresIrConfig(outer_tolerance=1e-8))
    assert st[0].converged and st[0].final_outer_residual <= 1e-8
    assert rel_l2(X[:, 0], np.linalg.solve(c1_block["A"], b)) < 1e-7


def test_lowp_update_after_reassembly(ctx, c1_block):
    """LowpMatrix.update re-converts the system's matrix after a new assemble."""
    sys = sa.TbemSystem(ctx, c1_block["mesh"].num_elements)
    ph = sa.PhysicsParams.from_wavenumber(1.0)
    c1_block["ws"].assemble(ph, ph.burton_miller_beta(), sys)
    lp = sa.LowpMatrix(sys, "fp16")
    D1 = lp.dequantized()
    ph2 = sa.PhysicsParams.from_wavenumber(3.0)
    c1_block["ws"].assemble(ph2, ph2.burton_miller_beta(), sys)
    lp.update()
    D2 = lp.dequantized()
    A2 = sys.matrix
    assert rel_fro(D2, A2) < U["fp16"]
    assert rel_fro(D1, A2) > 1e-2


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_lowp_copy_is_bit_exact(ctx, c1_block, fmt):
    """The device conversion of the assembled matrix equals the oracle's
    row-scaled storage bit for bit (halves, padding and row scales)."""
    from oracle import lowp as olp

    lp = sa.LowpMatrix(c1_block["sys"], fmt)
    halves, inv, ld = lp.download()
    ref_h, ref_inv = olp.row_scaled_halves(c1_block["A"], fmt, ld)
    assert np.array_equal(inv, ref_inv)
    bad = np.argwhere(halves != ref_h)
    for r, c in bad[:8]:
        a = c1_block["A"][r, c // 2]
        v = (a.real, a.imag)[c % 2] / inv[r]
        print(f"row {r} col {c}: scaled {v!r} device {halves[r, c]:#06x} oracle {ref_h[r, c]:#06x}")
    assert bad.shape[0] == 0, f"{bad.shape[0]} of {halves.size} halves differ"


@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_row_scales_from_assembly_equal_the_matrix_pass(ctx, fmt):
    """While a system has a low-precision copy, its value assembly leaves
    row maxima (far blocks without their near pairs and diagonal, near and
    diagonal entries folded in), and the next conversion takes its row scales
    from them instead of reading the matrix twice.  The first copy (made after
    an assembly that ran without a copy: the reading pass) equals the oracle's
    storage of the downloaded matrix bit for bit, rigid and mixed-BC meshes
    alike; a derivative assembly into the same system invalidates the maxima
    and the reading pass again matches the oracle; a new value assembly (now
    with maxima) gives the first copy again bit for bit, and so does the
    reading pass once the matrix's device address has been handed out."""
    from oracle import lowp as olp
    from sonium_amd.mesh import ellipsoid

    rng = np.random.default_rng(3)
    cases = [(icosphere(1.0, 3), 3.0, {}),
             (ellipsoid((1.3, 0.8, 0.6), 3), 5.0, {}),
             (icosphere(1.0, 3), 2.0, dict(bc_type=(rng.random(1280) < 0.3).astype(np.int32),
                                           bc_value=rng.standard_normal(1280) + 0j,
                                           admittance=np.where(rng.random(1280) < 0.3, 0.003 - 0.001j, np.nan)))]
    for m, k, bcs in cases:
        tm = sa.TbemMesh(ctx, m, **bcs)
        ws = sa.TbemAssemblyWorkspace(ctx, tm)
        ph = sa.PhysicsParams.from_wavenumber(k)
        beta = ph.burton_miller_beta()
        sys_ = ws.assemble(ph, beta)
        assert ws.stats["near_pairs"] > 0
        lp = sa.LowpMatrix(sys_, fmt)
        h1, inv1, ld = lp.download()
        ref_h, ref_inv = olp.row_scaled_halves(sys_.matrix, fmt, ld)
        assert np.array_equal(inv1, ref_inv) and np.array_equal(h1, ref_h)
        ws.assemble_wavenumber_derivative(ph, system=sys_)
        lp.update()
        h2, inv2, _ = lp.download()
        ref_h, ref_inv = olp.row_scaled_halves(sys_.matrix, fmt, ld)
        assert np.array_equal(inv2, ref_inv) and np.array_equal(h2, ref_h)
        ws.assemble(ph, beta, sys_)
        lp.update()
        h3, inv3, _ = lp.download()
        assert np.array_equal(inv3, inv1) and np.array_equal(h3, h1)
        sys_.device_pointers()
        lp.update()
        h4, inv4, _ = lp.download()
        assert np.array_equal(inv4, inv1) and np.array_equal(h4, h1)
        lp.close()
        sys_.close()
        ws.close()
        tm.close()


def test_device_block_solvers_equal_host_calls(ctx, c1_block):
    """gmres_block_device / gmres_ir_device on HBM-resident blocks give the host
    calls' solutions exactly; a zero start forms no FP64 product for the first
    residual and the final check reuses the last true residual."""
    op = sa.DenseMatrixOperator(c1_block["sys"])
    B = c1_block["B"]
    n, k = B.shape
    Bd, Xd = sa.DeviceVector(ctx, n * k), sa.DeviceVector(ctx, n * k)
    Bd.upload(B)
    cfg = sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    X, sols = sa.gmres_block(op, B, cfg)
    dsols = sa.gmres_block_device(op, Bd, k, Xd, cfg)
    assert np.array_equal(Xd.download().reshape(n, k), X)
    assert [s.iterations for s in dsols] == [s.iterations for s in sols]
    assert [s.residual for s in dsols] == [s.residual for s in sols]
    lp = sa.LowpMatrix(c1_block["sys"], "fp16")
    icfg = sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8)
    Xi, st = sa.gmres_ir(op, lp, B, icfg)
    dst = sa.gmres_ir_device(op, lp, Bd, k, Xd, icfg)
    assert np.array_equal(Xd.download().reshape(n, k), Xi)
    assert [s.final_outer_residual for s in dst] == [s.final_outer_residual for s in st]
    outer = max(s.outer_iterations for s in st)
    assert all(s.converged for s in st)
    # one block product per outer step plus one true-residual check per step at
    # which some column finished: none for the zero start, none for the final check
    finishes = len({s.outer_iterations for s in st})
    assert outer + 1 <= st[0].fp64_products <= outer + finishes
    with pytest.raises(ValueError):
        sa.gmres_ir_device(op, lp, Bd, k + 1, Xd, icfg)


@pytest.mark.parametrize("mesh_args", [("ico", 1), ("fib", (11, 7)), ("fib", (13, 13))])
@pytest.mark.parametrize("fmt", ["fp16", "bf16"])
def test_lowp_ragged_sizes(ctx, mesh_args, fmt):
    """Panel counts that are not a multiple of the 32-row tile (80, 140, 312):
    the packed copy downloads bit for bit as the oracle's row-major storage
    (padding rows never leak), and products for k on both sides of the
    32-column tile boundary equal the exactly quantized product."""
    from oracle import lowp as olp
    from sonium_amd.mesh import fibonacci_sphere

    kind, arg = mesh_args
    m = icosphere(1.0, arg) if kind == "ico" else fibonacci_sphere(1.0, *arg)
    n = m.num_elements
    assert n % 32
    tm = sa.TbemMesh(ctx, m)
    ph = sa.PhysicsParams.from_wavenumber(3.0)
    system = sa.TbemAssemblyWorkspace(ctx, tm).assemble(ph, ph.burton_miller_beta())
    A = system.matrix
    lp = sa.LowpMatrix(system, fmt)
    halves, inv, ld = lp.download()
    ref_h, ref_inv = olp.row_scaled_halves(A, fmt, ld)
    assert np.array_equal(inv, ref_inv) and np.array_equal(halves, ref_h)
    D = lp.dequantized()
    rng = np.random.default_rng(n)
    for k in (1, 3, 32, 33, 64):
        X = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
        Y = lp.apply(X if k > 1 else X[:, 0]).reshape(n, k)
        assert rel_fro(Y, D @ olp.quantized_columns(X, fmt)) < 5e-6, k


def test_block_gmres_initial_guess_matches_single(c1_block):
    """With X0 given, the first residual is formed from A X0 (no zero-start
    shortcut); every column then follows gmres_with_guess."""
    op = sa.DenseMatrixOperator(c1_block["sys"])
    B = c1_block["B"][:, :4]
    A = c1_block["A"]
    X0 = 0.9 * np.linalg.solve(A, B) + 0.01
    cfg = sa.GmresConfig(max_iterations=100, restart=50, tolerance=1e-8)
    X, sols = sa.gmres_block(op, B, cfg, X0=X0)
    for c in range(B.shape[1]):
        single = sa.gmres_with_guess(op, B[:, c], X0[:, c], cfg)
        assert sols[c].converged and sols[c].residual <= 1e-8
        assert sols[c].iterations == single.iterations
        assert rel_l2(X[:, c], single.x) < 1e-10
    # an exact guess converges without an Arnoldi step
    Xe = np.linalg.solve(A, B)
    _, sols = sa.gmres_block(op, B, cfg, X0=Xe)
    assert all(s.converged and s.iterations == 0 for s in sols)
