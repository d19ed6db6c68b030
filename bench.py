"""Benchmark: solved right-hand sides per second at relative residual 1e-8.

One step = one pass of the hot path for one right-hand side on one GPU:
assemble the dense Burton–Miller matrix for the wavenumber (far, near and
self kernels), add the plane-wave right-hand side, and run GMRES to a
recomputed relative residual of 1e-8 — all with the mesh and its
frequency-invariant near-pair routes already resident in HBM (the
workspace is built once per mesh, as the reference's resident CUDA workspace
is; its build time is reported as ``setup_ms``).

Default workload: rigid sphere, icosphere with 20480 panels, k = 10,
beta = i/k, one plane wave along +z, GMRES(100), rtol 1e-8, FP64.

Printed: one JSON line with the metric, the per-kernel roofline of the
dominant kernel (timed with HIP events on the library's stream inside the
timed region; the events come from a per-context pool, and the plain and
mixed-precision legs are timed the same way), the chosen form against the
plain FP64 path for every configuration (the 40960-panel 64-RHS solve, the
256-ellipsoid batch and the frequency sweep are appended), and the CPU
baseline (the oracle's C++ restatement of the reference algorithm on this
host's cores).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import sonium_amd as sa  # noqa: E402
from sonium_amd.mesh import fibonacci_directions, fibonacci_sphere, icosphere  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FLOPS_PER_EVAL = 70      # FP64 flops per kernel evaluation (FMA = 2, sqrt/div/sincos = 1)
FLOPS_PER_COEFF = 14     # coefficient H*gamma*tau + E*beta per pair
FIELD_FLOPS_PER_QP = 47  # field kernel, rigid panels: FP64 flops per (point, panel, quadrature point)

# C5 batched sweep: solved points per side in the minimal-residual guesses
C5_NEIGHBOURS_PER_SIDE = 1

# HBM bytes per launch (read + written) measured with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE on this workload, scaled by a 1 GiB copy (FETCH_SIZE x 2 KiB,
# WRITE_SIZE x 1 KiB); profiles/zgemv_n_kernel_and_far_kernel_c2_icosphere20480_k10.md.
MEASURED_TRAFFIC = {
    ("zgemv_n_kernel", 20480): 6.7136e9 + 3.29e5,
    ("far_kernel", 20480): 1.309e8 + 6.7643e9,
}

WORKLOADS = {
    "c2": {"subdiv": 5, "k": 10.0, "restart": 100,
           "name": "rigid sphere icosphere 20480 panels, k=10, beta=i/k, 1 plane wave +z, GMRES(100) rtol 1e-8, complex128"},
    "c3": {"name": "many right-hand sides (separate report)"},
    "c4": {"name": "batch of 256 seeded rigid ellipsoids (separate report)"},
    "c5": {"name": "frequency sweep with learned warm start (separate report)"},
    "c1": {"subdiv": 3, "k": 2.0, "restart": 50,
           "name": "rigid sphere icosphere 1280 panels, k=2, beta=i/k, 1 plane wave +z, GMRES(50) rtol 1e-8, complex128"},
}


def host_info():
    """CPU model, logical CPUs, the CPUs this process may run on, and the
    thread count used: the affinity set, capped by OMP_NUM_THREADS when the
    host sets it (the GPU hosts allot each GPU a share of their cores that
    way)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    threads = affinity
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    return {"cpu_model": model, "nproc": nproc, "affinity_cpus": affinity, "threads_used": threads,
            "omp_num_threads": omp or None}


def cpu_baseline(m, k, beta, spmv_count, rows, threads):
    """The oracle (C++ restatement of the reference CPU path) on this host.

    value: the C2 step (20480 panels) from the assembly of rows [0, rows) and
    three matvecs of that row block, extrapolated to N (the full matrix is
    6.7 GB and its assembly minutes of CPU time).  Beside it: the 1280-panel
    solve timed in full, and the reference's own crossover measurement
    (bem_assembly_crossover.rs:128-152: build_tbem_system_stabilized + 10
    dense matvecs, icosphere, ka = 1; published 6.20 s at N = 1280 and
    83.50 s at N = 5120, bem_solver.rs:222-231) restated at both sizes."""
    import oracle

    info = host_info()
    n = m.num_elements
    rows = min(rows, n)
    t0 = time.perf_counter()
    A_rows, _ = oracle.build_tbem_system(m.nodes, m.triangles, k, beta, rows=(0, rows), threads=threads)
    t_asm = time.perf_counter() - t0
    x = np.ones(n, dtype=np.complex128)
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.zgemv(A_rows, x, threads=threads)
    t_mv = (time.perf_counter() - t0) / reps
    scale = n / rows
    t_solve = spmv_count * t_mv * scale
    t_total = t_asm * scale + t_solve
    del A_rows
    # 1280 panels, k = 2, GMRES(50): timed in full
    m1 = icosphere(1.0, 3)
    t0 = time.perf_counter()
    A1, _ = oracle.build_tbem_system(m1.nodes, m1.triangles, 2.0, 0.5j, threads=threads)
    c1, n1, _ = m1.element_geometry()
    b1 = oracle.plane_wave_rhs(c1, n1, 2.0, 0.5j)
    t1a = time.perf_counter() - t0
    x1, st1 = oracle.gmres(A1, b1, restart=50, max_iterations=100, tol=1e-8, threads=threads)
    t1 = time.perf_counter() - t0
    # the reference's crossover sizes (ka = 1 on the unit sphere)
    cross = {}
    for subdiv, th, published in ((3, 1, 6.20), (3, threads, 6.20), (4, threads, 83.50)):
        mm = icosphere(1.0, subdiv)
        t0 = time.perf_counter()
        Am, _ = oracle.build_tbem_system(mm.nodes, mm.triangles, 1.0, 1.0j, threads=th)
        ta = time.perf_counter() - t0
        xm = np.ones(mm.num_elements, dtype=np.complex128)
        t0 = time.perf_counter()
        for _ in range(10):
            oracle.zgemv(Am, xm, threads=th)
        tmv = time.perf_counter() - t0
        cross[f"n{mm.num_elements}_threads{th}"] = {"assembly_s": ta, "matvec10_s": tmv, "total_s": ta + tmv,
                                                     "reference_published_s": published,
                                                     "reference_over_restatement": published / (ta + tmv)}
        del Am
    return {
        "value": 1.0 / t_total,
        "unit": "RHS/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle assembly of rows [0,{rows}) of {n} and {reps} matvecs of that row block, "
                   f"extrapolated x{scale:.1f}; solve = {spmv_count} matvecs (the GPU run's count); "
                   f"assembly {t_asm * scale:.1f} s + solve {t_solve:.1f} s per RHS"),
        "host": info,
        "oracle_library": os.path.basename(oracle.library_path()),
        "c1_full": {"panels": m1.num_elements, "assembly_s": t1a, "total_s": t1, "iterations": st1["iterations"],
                    "rhs_per_s": 1.0 / t1},
        "reference_crossover": cross,
    }


def run_ir_leg(ctx, ws, tm, ph, beta, n, steps, fmt, inner_tol):
    """Mixed-precision form of the same step: assemble FP64, convert to the
    row-scaled complex FP16/BF16 copy, FGMRES-IR to an FP64-recomputed
    residual of 1e-8.  Timed like the plain leg: kernel events on (pooled)."""
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    inc = sa.IncidentField.plane_wave_z()
    cfg = sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8, max_inner_iterations=50,
                           inner_tolerance=inner_tol)
    ws.assemble(ph, beta, system)
    lp = sa.LowpMatrix(system, fmt)
    bd, xd = sa.DeviceVector(ctx, n), sa.DeviceVector(ctx, n)

    def step():
        ws.assemble(ph, beta, system)
        lp.update()
        inc.write_device(bd, tm, ph, beta)
        return sa.gmres_ir_device(op, lp, bd, 1, xd, cfg)

    step()
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    outs = [step() for _ in range(steps)]
    ctx.synchronize()
    el = time.perf_counter() - t0
    stats = {name: ctx.kernel_stats(name) for name in ("lowp", "matvec", "krylov_vector", "assemble_far")}
    ctx.set_profiling(False)
    st = outs[-1]
    x = xd.download()
    if not all(o[0].converged and o[0].final_outer_residual <= 1e-8 for o in outs):
        raise SystemExit(f"bench: mixed-precision solve did not converge: {st[0]}")
    return {"form": f"{fmt} storage + FP32 accumulate, FGMRES-IR with FP64 residual/Arnoldi", "value": steps / el,
            "ms_per_step": el / steps * 1e3, "outer_iterations": st[0].outer_iterations,
            "inner_iterations": st[0].total_inner_iterations, "residual": st[0].final_outer_residual,
            "fp64_products": st[0].fp64_products, "lowp_products": st[0].lowp_products,
            "kernel_ms_per_step": {k: v[1] / steps for k, v in stats.items()}}, x


def run_c3(args, ctx=None):
    """Many right-hand sides: 40960-panel Fibonacci sphere, k = 14, 64 plane waves.

    plain: the reference's multi-RHS solver (solve_many -> block GMRES with one
    shared block Krylov basis, block_gmres.rs:666-1120) in FP64, restart 100.
    Also timed: the column-by-column GMRES (each column the single-RHS
    method, one block product per step).  Chosen: FP16 and BF16 storage with
    FGMRES iterative refinement (FP64 residual, Arnoldi and Givens).
    """
    ctx = ctx or sa.Context(0)
    m = fibonacci_sphere(1.0, 129, 160)
    n = m.num_elements
    k = 14.0
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    dirs = fibonacci_directions(64)
    inc = sa.IncidentField(dirs)
    # the low-precision copies exist before the assembly, as in a solver that
    # keeps them: the assembly then leaves the row maxima their conversion reads
    lps = {fmt: sa.LowpMatrix(system, fmt) for fmt in ("fp16", "bf16")}
    t0 = time.perf_counter()
    ws.assemble(ph, beta, system)
    ctx.synchronize()
    t_asm = time.perf_counter() - t0
    # right-hand sides and solutions stay in HBM (row-major n x 64 blocks)
    B = sa.DeviceVector(ctx, n * 64)
    Xd = sa.DeviceVector(ctx, n * 64)
    inc.write_device(B, tm, ph, beta)
    res = {}
    ctx.set_profiling(True)
    ctx.synchronize()
    ctx.reset_stats()
    t0 = time.perf_counter()
    bsol = sa.block_gmres_device(op, B, 64, Xd, sa.BlockGmresConfig(max_iterations=100, restart=100, tolerance=1e-8))
    ctx.synchronize()
    t_block = time.perf_counter() - t0
    X64 = Xd.download().reshape(n, 64)
    blk = ctx.kernel_stats("block")
    kry = ctx.kernel_stats("krylov_vector")
    if not (bsol.converged and bsol.true_residuals.max() <= 1e-8):
        raise SystemExit(f"bench c3: block GMRES did not converge: {bsol.true_residuals.max()}")
    res["plain_fp64"] = {"solver": "block GMRES, shared basis (solve_many)", "solve_s": t_block,
                         "rhs_per_s_excl_assembly": 64 / t_block, "rhs_per_s_incl_assembly": 64 / (t_block + t_asm),
                         "block_steps": bsol.iterations, "basis_vectors": bsol.max_basis,
                         "column_products": bsol.spmv_count, "max_residual": float(bsol.true_residuals.max()),
                         "fp64_block_product_avg_ms": blk[1] / max(blk[0], 1), "fp64_block_products": blk[0],
                         "block_krylov_ms": kry[1]}
    ctx.reset_stats()
    X2 = sa.DeviceVector(ctx, n * 64)
    t0 = time.perf_counter()
    sols = sa.gmres_block_device(op, B, 64, X2, sa.GmresConfig(max_iterations=100, restart=100, tolerance=1e-8))
    ctx.synchronize()
    t_cols = time.perf_counter() - t0
    Xc = X2.download().reshape(n, 64)
    X2.close()
    res["per_column_fp64"] = {"solver": "GMRES per column, one block product per step", "solve_s": t_cols,
                              "rhs_per_s_excl_assembly": 64 / t_cols,
                              "max_iterations": max(s.iterations for s in sols),
                              "max_residual": max(s.residual for s in sols),
                              "max_diff_vs_block": float(np.max(np.linalg.norm(Xc - X64, axis=0)
                                                                / np.linalg.norm(X64, axis=0)))}
    for fmt, itol in (("fp16", 1e-3), ("bf16", 1e-2)):
        # the low-precision copy is allocated once, like the FP64 system; the
        # timed conversion refills it from the current FP64 matrix
        lp = lps[fmt]
        ctx.synchronize()
        ctx.reset_stats()
        t0 = time.perf_counter()
        lp.update()
        ctx.synchronize()
        t_conv = time.perf_counter() - t0
        t0 = time.perf_counter()
        st = sa.gmres_ir_device(op, lp, B, 64, Xd, sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8,
                                                                     max_inner_iterations=50, inner_tolerance=itol))
        ctx.synchronize()
        t_ir = time.perf_counter() - t0
        X = Xd.download().reshape(n, 64)
        lpk = ctx.kernel_stats("lowp")
        blk = ctx.kernel_stats("block")
        err = np.linalg.norm(X - X64, axis=0) / np.linalg.norm(X64, axis=0)
        lp_avg = lpk[1] / max(lpk[0], 1) / 1e3
        lp_bytes = 4 * n * n
        lp_flops = 8 * n * n * 64
        res[fmt + "_ir"] = {"solve_s": t_ir, "convert_s": t_conv,
                            "rhs_per_s_excl_assembly": 64 / (t_ir + t_conv),
                            "rhs_per_s_incl_assembly": 64 / (t_ir + t_conv + t_asm),
                            "max_outer": max(s.outer_iterations for s in st),
                            "max_inner": max(s.total_inner_iterations for s in st),
                            "max_residual": max(s.final_outer_residual for s in st),
                            "error_vs_fp64_per_rhs_max": float(err.max()),
                            "error_vs_fp64_per_rhs_median": float(np.median(err)),
                            "converged": all(s.converged for s in st),
                            "lowp_gemm_avg_ms": lp_avg * 1e3, "lowp_gemm_gbs": lp_bytes / lp_avg / 1e9,
                            "lowp_gemm_tflops": lp_flops / lp_avg / 1e12,
                            "fp64_block_product_avg_ms": blk[1] / max(blk[0], 1)}
        res[fmt + "_ir"]["vs_plain_excl_assembly"] = (res[fmt + "_ir"]["rhs_per_s_excl_assembly"]
                                                      / res["plain_fp64"]["rhs_per_s_excl_assembly"])
        res[fmt + "_ir"]["vs_plain_incl_assembly"] = (res[fmt + "_ir"]["rhs_per_s_incl_assembly"]
                                                      / res["plain_fp64"]["rhs_per_s_incl_assembly"])
        lp.close()
    ctx.set_profiling(False)
    lp = res["fp16_ir"]
    zg_ms = res["plain_fp64"]["fp64_block_product_avg_ms"]
    return {"metric": "solved RHS/s at rel. residual 1e-8 (multi-RHS)",
            "workload": "C3: Fibonacci sphere 40960 panels, k=14, 64 plane waves",
            "panels": n, "assembly_s": t_asm, "results": res,
            "roofline_lp_gemm": {"kernel": "lp_gemm_kernel", "bound": "hbm", "achieved": lp["lowp_gemm_gbs"],
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": lp["lowp_gemm_gbs"] / HBM_PEAK_GBS},
            "roofline_fp64_block_product": {"kernel": "zgemm_f64_kernel", "bound": "fp64-mfma",
                                            "achieved": 8 * n * n * 64 / (zg_ms * 1e-3) / 1e12,
                                            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                            "frac": 8 * n * n * 64 / (zg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}}


def _percentiles(values):
    v = np.asarray(values, dtype=np.float64)
    return {"p50_ms": float(np.percentile(v, 50)), "p99_ms": float(np.percentile(v, 99)),
            "max_ms": float(v.max())}


def run_c4(args, ctx=None):
    """256 independent rigid ellipsoids (320-5120 panels, k U[1,8], 1-16 plane waves).

    plain: the single-system FP64 path instance by instance, as a caller of
    the reference-shaped API runs it: assemble into a device system, write the
    plane-wave right-hand sides on the device, solve all of an instance's
    right-hand sides with the reference's multi-RHS solver (block GMRES,
    shared basis, restart 100).  Systems and right-hand-side/solution blocks
    are allocated once per size before the timed loop.  The same loop with the
    column-by-column GMRES is timed too.  Per-instance latency = wall time.
    batched: every instance in one batch: persistent assembly, then lock-step
    GMRES (FP64) or FP16 iterative refinement with each cycle replayed from a
    captured HIP graph; per-instance latency = device clock from the start of
    the batch assembly to the instance's last converged column.
    """
    from sonium_amd.mesh import random_ellipsoid_batch

    ctx = ctx or sa.Context(0)
    specs = random_ellipsoid_batch(256, seed=20261202)
    tms = [sa.TbemMesh(ctx, s["mesh"]) for s in specs]
    ncols = np.array([len(s["directions"]) for s in specs])
    total_rhs = int(ncols.sum())
    sizes = np.array([s["mesh"].num_elements for s in specs])
    print(f"c4: {len(specs)} instances, {total_rhs} right-hand sides, panels {sizes.sum()}", file=sys.stderr, flush=True)

    # plain FP64, one instance after another (routes built beforehand, as setup)
    wss = [sa.TbemAssemblyWorkspace(ctx, tm) for tm in tms]
    systems = {int(nn): sa.TbemSystem(ctx, int(nn)) for nn in np.unique(sizes)}
    ops = {nn: sa.DenseMatrixOperator(sy) for nn, sy in systems.items()}
    blocks = {}
    for nn, c in set(zip(sizes.tolist(), ncols.tolist())):
        blocks[(nn, c)] = (sa.DeviceVector(ctx, nn * c), sa.DeviceVector(ctx, nn * c))
    bcfg = sa.BlockGmresConfig(max_iterations=10, restart=100, tolerance=1e-8)
    ccfg = sa.GmresConfig(max_iterations=10, restart=100, tolerance=1e-8)
    plain = {}
    plain_x = {}
    for form in ("block", "per_column"):
        lat, iters = [], []
        ctx.synchronize()
        t_all = time.perf_counter()
        for i, (s, tm, ws) in enumerate(zip(specs, tms, wss)):
            t0 = time.perf_counter()
            nn, c = int(sizes[i]), int(ncols[i])
            ph = sa.PhysicsParams.from_wavenumber(s["k"])
            beta = ph.burton_miller_beta()
            ws.assemble(ph, beta, systems[nn])
            Bd, Xd = blocks[(nn, c)]
            sa.IncidentField(s["directions"]).write_device(Bd, tm, ph, beta)
            if form == "block":
                sol = sa.block_gmres_device(ops[nn], Bd, c, Xd, bcfg)
                ok = sol.converged and sol.true_residuals.max() <= 1e-8
                iters.append(sol.iterations)
            else:
                sols = sa.gmres_block_device(ops[nn], Bd, c, Xd, ccfg)
                ok = all(q.converged and q.residual <= 1e-8 for q in sols)
                iters.append(max(q.iterations for q in sols))
            ctx.synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
            if not ok:
                raise SystemExit(f"bench c4: plain {form} solve of instance {i} did not converge")
            if form == "block" and i % 8 == 0:
                plain_x[i] = Xd.download().reshape(nn, c)
        el = time.perf_counter() - t_all
        plain[form] = {"seconds": el, "rhs_per_s": total_rhs / el, **_percentiles(lat),
                       "max_iterations": int(max(iters)), "mean_iterations": float(np.mean(iters))}
        print(f"c4: plain {form} {el:.2f} s", file=sys.stderr, flush=True)
    for Bd, Xd in blocks.values():
        Bd.close()
        Xd.close()
    for o in list(ops.values()) + list(systems.values()):
        o.close()
    del wss
    t_plain = plain["block"]["seconds"]

    t0 = time.perf_counter()
    batch = sa.Batch(ctx, tms)
    batch.set_problem([s["k"] for s in specs], [s["directions"] for s in specs])
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    best, table = batch.autotune()
    tune_s = time.perf_counter() - t0
    print(f"c4: batch setup {setup_s:.2f} s, autotune {tune_s:.2f} s -> {best}", file=sys.stderr, flush=True)
    forms = {}
    for form, fmt in (("batched_fp64", "fp64"), ("batched_fp16_ir", "fp16_ir")):
        for rep in range(2):  # first run captures the cycle graph
            ctx.synchronize()
            t0 = time.perf_counter()
            batch.assemble()
            if fmt != "fp64":
                batch.convert("fp16")
            if fmt == "fp64":
                sts = batch.solve("fp64", restart=100, max_cycles=10, tolerance=1e-8, use_graph=not args.no_graph)
            else:
                sts = batch.solve("fp16_ir", restart=30, max_cycles=30, tolerance=1e-8, inner_tolerance=1e-3,
                                  use_graph=not args.no_graph)
            ctx.synchronize()
            el = time.perf_counter() - t0
        if not all(q.converged for q in sts):
            raise SystemExit(f"bench c4: {form} left columns unconverged")
        lat, col, errs = [], 0, []
        for i, s in enumerate(specs):
            c = int(ncols[i])
            lat.append(max(q.completion_ms for q in sts[col: col + c]))
            col += c
        for i in sorted(plain_x):
            X = batch.solution(i)
            errs.append(np.linalg.norm(X - plain_x[i], axis=0) / np.linalg.norm(plain_x[i], axis=0))
        errs = np.concatenate(errs)
        forms[form] = {"seconds": el, "rhs_per_s": total_rhs / el, **_percentiles(lat),
                       "max_iterations": max(q.iterations for q in sts), "max_cycles": max(q.cycles for q in sts),
                       "max_residual": max(q.residual for q in sts),
                       "error_vs_plain_per_rhs_max": float(errs.max()),
                       "error_vs_plain_per_rhs_median": float(np.median(errs)),
                       "vs_plain": (total_rhs / el) / (total_rhs / t_plain)}
        print(f"c4: {form} {el:.3f} s", file=sys.stderr, flush=True)
    batch.close()
    # Latency knob: the same solve over sub-batches served one after another
    # (each sub-batch's instances assembled and solved together); an
    # instance's service latency is counted from its sub-batch's start, and
    # its job latency from the start of the whole job.  "size_classes" serves
    # one sub-batch per panel count, smallest first, so the many cheap
    # instances finish before the large ones start.
    sub = {}
    order_by_size = sorted(range(len(specs)), key=lambda i: (sizes[i], i))
    for size in (64, 16, "size_classes"):
        if size == "size_classes":
            groups = [[i for i in order_by_size if sizes[i] == nn] for nn in sorted(set(sizes.tolist()))]
        else:
            groups = [list(range(g, min(g + size, len(specs)))) for g in range(0, len(specs), size)]
        batches = []
        for idx in groups:
            bb = sa.Batch(ctx, [tms[i] for i in idx])
            bb.set_problem([specs[i]["k"] for i in idx], [specs[i]["directions"] for i in idx])
            bb.set_tuning(**best)
            batches.append(bb)
        for rep in range(2):  # the first pass captures each sub-batch's cycle graph
            lat, job_lat = [], []
            ctx.synchronize()
            t_all = time.perf_counter()
            for idx, bb in zip(groups, batches):
                t0 = time.perf_counter()
                offset_ms = (t0 - t_all) * 1e3
                bb.assemble()
                bb.convert("fp16")
                sts = bb.solve("fp16_ir", restart=30, max_cycles=30, tolerance=1e-8, inner_tolerance=1e-3,
                               use_graph=not args.no_graph)
                if not all(q.converged for q in sts):
                    raise SystemExit("bench c4: a sub-batch left columns unconverged")
                col = 0
                for i in idx:
                    c = int(ncols[i])
                    lat.append(max(q.completion_ms for q in sts[col: col + c]))
                    job_lat.append(offset_ms + lat[-1])
                    col += c
            ctx.synchronize()
            el = time.perf_counter() - t_all
        jp = _percentiles(job_lat)
        key = "size_classes_smallest_first" if size == "size_classes" else f"sub_batches_of_{size}"
        sub[key] = {"seconds": el, "rhs_per_s": total_rhs / el, **_percentiles(lat),
                    "job_p50_ms": jp["p50_ms"], "job_p99_ms": jp["p99_ms"],
                    "vs_plain": (total_rhs / el) / (total_rhs / t_plain)}
        for bb in batches:
            bb.close()
        print(f"c4: {key}: {el:.3f} s, job p50 {jp['p50_ms']:.1f} ms", file=sys.stderr, flush=True)
    forms["batched_fp16_ir_sub_batches"] = sub
    line = {"metric": "solved RHS/s at rel. residual 1e-8 (batch)",
            "workload": "C4: 256 rigid ellipsoids, semi-axes U[0.5,1.5], 320/1280/5120 panels, k U[1,8], 1-16 plane waves, seed 20261202",
            "instances": len(specs), "rhs": total_rhs, "panels": int(sizes.sum()),
            "matrix_gb_fp64": float((sizes.astype(np.float64) ** 2).sum() * 16 / 1e9),
            "plain_fp64_sequential": plain["block"], "plain_fp64_sequential_per_column": plain["per_column"],
            "batched": forms, "autotune": {"best": best, "candidates": table, "seconds": tune_s},
            "batch_setup_s": setup_s}
    return line


def run_c5(args, ctx=None):
    """Frequency sweep with the learned warm start.

    Training data: the 10240-panel Fibonacci sphere at 512 wavenumbers in
    [1, 12] (+z incidence) and 16 seeded ellipsoid variants of it (semi-axes
    U[0.75, 1.25], 93 wavenumbers each, random incidence): 2000 solver
    solutions.  Evaluation: 256 held-out wavenumbers (midpoints of the
    training grid) on the sphere; per wavenumber a new matrix is assembled and
    solved to 1e-8 from (a) zero, (b) the previous wavenumber's solution,
    (c) the network's guess, (d) GCRO-DR recycling, alone and with (b).
    Batched forms: 4 interleaved groups of 64 frequencies (indices g, g+4,
    ...; order 0, 2, 1, 3), each group one batch on the shared mesh (its
    equally spaced wavenumbers take the frequency-family assembly), solved by
    lock-step FP16 refinement to the FP64 residual bar from zero, previous,
    learned, or minimal-residual neighbour-combination guesses.
    """
    from sonium_amd import warmstart as wsm
    from sonium_amd.mesh import Mesh, uniform_directions

    ctx = ctx or sa.Context(0)
    base = fibonacci_sphere(1.0, 65, 80)
    n = base.num_elements
    cfg = sa.GmresConfig(max_iterations=20, restart=100, tolerance=1e-8)
    rng = np.random.Generator(np.random.PCG64(20261203))
    meshes = [("sphere", base, np.linspace(1.0, 12.0, 512), None)]
    for v in range(16):
        axes = rng.uniform(0.75, 1.25, size=3)
        ks = np.sort(rng.uniform(1.0, 12.0, size=93))
        meshes.append((f"ellipsoid{v}", Mesh(base.nodes * axes[None, :], base.triangles), ks,
                       uniform_directions(rng, 93)))
    inputs, targets = [], []
    t0 = time.perf_counter()
    gen_iters = []
    for name, m, ks, dirs in meshes:
        tm = sa.TbemMesh(ctx, m)
        c, nn, _ = tm.geometry()
        ws = sa.TbemAssemblyWorkspace(ctx, tm)
        system = sa.TbemSystem(ctx, n)
        op = sa.DenseMatrixOperator(system)
        for j, k in enumerate(ks):
            d = np.array([0.0, 0.0, 1.0]) if dirs is None else dirs[j]
            ph = sa.PhysicsParams.from_wavenumber(float(k))
            beta = ph.burton_miller_beta()
            ws.assemble(ph, beta, system)
            b = sa.IncidentField([d]).compute_rhs_with_beta(tm, ph, beta)
            sol = sa.gmres(op, b, cfg)
            if not sol.converged:
                raise SystemExit(f"bench c5: training solve {name} k={k} did not converge")
            gen_iters.append(sol.iterations)
            inputs.append(wsm.panel_inputs(c, nn, float(k), d))
            targets.append(wsm.density_to_target(sol.x, c, float(k), d))
        print(f"c5: generated {name} ({len(ks)} solves)", file=sys.stderr, flush=True)
    gen_s = time.perf_counter() - t0
    n_solutions = len(inputs)
    net = wsm.WarmStartMLP(ctx, width=512, seed=20261204)
    net.set_dataset(np.concatenate(inputs), np.concatenate(targets))
    del inputs, targets
    steps, batch_size = 16000, 16384
    ctx.synchronize()
    t0 = time.perf_counter()
    losses = net.train(steps, batch=batch_size, lr_max=2e-3, lr_min=2e-5, seed=7)
    ctx.synchronize()
    train_s = time.perf_counter() - t0
    print(f"c5: trained {steps} steps in {train_s:.1f} s, loss {losses[:3]} -> {losses[-3:]}", file=sys.stderr,
          flush=True)

    # evaluation on held-out wavenumbers
    tm = sa.TbemMesh(ctx, base)
    c, nn, _ = tm.geometry()
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    kgrid = np.linspace(1.0, 12.0, 512)
    ks = 0.5 * (kgrid[0::2] + kgrid[1::2])
    d = np.array([0.0, 0.0, 1.0])
    res = {}
    x_zero = {}
    for mode in ("zero", "previous", "learned"):
        iters, used = [], 0
        prev = None
        guess_err, guess_res = [], []
        ctx.synchronize()
        t0 = time.perf_counter()
        for k in ks:
            ph = sa.PhysicsParams.from_wavenumber(float(k))
            beta = ph.burton_miller_beta()
            ws.assemble(ph, beta, system)
            b = sa.IncidentField([d]).compute_rhs_with_beta(tm, ph, beta)
            x0 = None
            if mode == "previous":
                x0 = prev
            elif mode == "learned":
                x0 = net.guess(c, nn, float(k), d)
            sol, u, r0 = wsm.solve_with_guess(op, b, x0, cfg, with_residual=True)
            used += int(u)
            if not (sol.converged and sol.residual <= 1e-8):
                raise SystemExit(f"bench c5: {mode} solve at k={k} did not converge")
            iters.append(sol.iterations)
            prev = sol.x
            if x0 is not None:
                guess_res.append(r0)
            if mode == "zero":
                x_zero[float(k)] = sol.x
            elif x0 is not None:
                ref = x_zero[float(k)]
                guess_err.append(float(np.linalg.norm(x0 - ref) / np.linalg.norm(ref)))
        ctx.synchronize()
        el = time.perf_counter() - t0
        res[mode] = {"rhs_per_s": len(ks) / el, "seconds": el, "mean_iterations": float(np.mean(iters)),
                     "max_iterations": int(np.max(iters)), "guesses_used": used}
        if guess_err:
            res[mode]["guess_rel_error_median"] = float(np.median(guess_err))
            res[mode]["guess_rel_error_max"] = float(np.max(guess_err))
            res[mode]["guess_rel_residual_median"] = float(np.median(guess_res))
        print(f"c5: {mode} {el:.2f} s, mean iterations {np.mean(iters):.1f}", file=sys.stderr, flush=True)
    # the reference's other sweep accelerator: GCRO-DR recycling 10 harmonic Ritz vectors from
    # one frequency to the next (frequency_sweep.rs:495-508,636-650), alone and with the
    # previous solution as the initial guess; each solve also spends 10 products rebuilding C
    gcfg = sa.GcroDrConfig.from_gmres(cfg)
    for mode in ("gcro_dr", "gcro_dr_previous"):
        sub = sa.GcroDrSubspace(ctx, n, 10)
        iters, spmv = [], []
        prev = None
        ctx.synchronize()
        t0 = time.perf_counter()
        for k in ks:
            ph = sa.PhysicsParams.from_wavenumber(float(k))
            beta = ph.burton_miller_beta()
            ws.assemble(ph, beta, system)
            b = sa.IncidentField([d]).compute_rhs_with_beta(tm, ph, beta)
            sol = sa.gcro_dr(op, b, prev if mode == "gcro_dr_previous" else None, sub, gcfg)
            if not (sol.converged and sol.residual <= 1e-8):
                raise SystemExit(f"bench c5: {mode} solve at k={k} did not converge")
            iters.append(sol.iterations)
            spmv.append(sol.spmv_count)
            prev = sol.x
        ctx.synchronize()
        el = time.perf_counter() - t0
        res[mode] = {"rhs_per_s": len(ks) / el, "seconds": el, "mean_iterations": float(np.mean(iters)),
                     "max_iterations": int(np.max(iters)), "mean_products": float(np.mean(spmv)),
                     "recycled_vectors": 10}
        sub.close()
        print(f"c5: {mode} {el:.2f} s, mean iterations {np.mean(iters):.1f}, mean products {np.mean(spmv):.1f}",
              file=sys.stderr, flush=True)
    # batched forms (sonium_amd.sweep): 4 interleaved groups of 64 frequencies, coarse to fine,
    # each one frequency-family batch solved by lock-step FP16 refinement to the FP64 residual bar
    from sonium_amd import sweep
    stride = 4
    t0 = time.perf_counter()
    sb = sweep.SweepBatch(ctx, base, len(ks) // stride)
    sb.ensure_store(len(ks))
    batch_setup_s = time.perf_counter() - t0
    # the network's inference workspace for one whole group, allocated before the timed sweeps
    net.predict(np.zeros((len(ks) // stride * n, wsm.RAW), dtype=np.float32))
    x_ref = np.stack([x_zero[float(k)] for k in ks])
    batched_modes = {
        # name: (guess of the first group, guess of later groups)
        "batched_fp16_ir_zero": (None, None),
        "batched_fp16_ir_previous": (None, "previous"),
        "batched_fp16_ir_learned": ("learned", "learned"),
        "batched_fp16_ir_neighbours": (None, "neighbours"),
        "chosen_batched_fp16_ir_learned_then_neighbours": ("learned", "neighbours"),
        "batched_fp16_ir_learned_then_neighbours_and_learned": ("learned", "neighbours+learned"),
    }
    for mode, (first, later) in batched_modes.items():
        r = sweep.batched_sweep(ctx, base, ks, d, stride=stride, net=net, first_guess=first, later_guess=later,
                                per_side=C5_NEIGHBOURS_PER_SIDE, sweep_batch=sb, use_graph=not args.no_graph)
        errs = [float(np.linalg.norm(r.solutions[j] - x_ref[j]) / np.linalg.norm(x_ref[j]))
                for j in range(0, len(ks), 16)]
        res[mode] = {"rhs_per_s": len(ks) / r.seconds, "seconds": r.seconds,
                     "mean_iterations": float(r.iterations.mean()), "max_iterations": int(r.iterations.max()),
                     "max_residual": float(r.residuals.max()), "group": len(ks) // stride,
                     "batch_setup_s": batch_setup_s, "breakdown_s": r.breakdown_s,
                     "error_vs_fp64_max_sampled": float(np.max(errs))}
        if r.guess_residuals:
            res[mode]["guess_rel_residual_median"] = float(np.median(list(r.guess_residuals.values())))
        print(f"c5: {mode} {r.seconds:.2f} s, mean Arnoldi steps {r.iterations.mean():.1f}", file=sys.stderr,
              flush=True)
    for mode in ["previous", "learned", "gcro_dr", "gcro_dr_previous"] + list(batched_modes):
        res[mode]["vs_zero"] = res[mode]["rhs_per_s"] / res["zero"]["rhs_per_s"]
    for mode in batched_modes:
        res[mode]["vs_batched_zero"] = res[mode]["rhs_per_s"] / res["batched_fp16_ir_zero"]["rhs_per_s"]
    sb.close()
    net.close()
    line = {"metric": "solved RHS/s at rel. residual 1e-8 (frequency sweep)",
            "workload": "C5: Fibonacci sphere 10240 panels, held-out k in [1,12] (256), +z; MLP 4x512 BF16 trained on "
                        f"{n_solutions} solver solutions (sphere sweep + 16 ellipsoid variants)",
            "panels": n, "training": {"solutions": n_solutions, "samples": n_solutions * n, "generate_s": gen_s,
                                      "generation_mean_iterations": float(np.mean(gen_iters)),
                                      "steps": steps, "batch": batch_size, "train_s": train_s,
                                      "loss_first": float(losses[0]), "loss_last": float(np.mean(losses[-50:]))},
            "results": res}
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--cpu-rows", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-chosen-form", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the multi-RHS (C3) and batch (C4) comparisons that the default run appends")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched solves launch every kernel directly instead of replaying a captured cycle")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 3 plain device copies of BYTES bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        ctx = sa.Context(0)
        sa._check(sa._lib().sbh_debug_device_copy(ctx.handle, args.calibrate_copy, 3))
        print(json.dumps({"copy_bytes": args.calibrate_copy, "launches": 3}))
        return
    if args.workload in ("c3", "c4", "c5"):
        print(json.dumps({"c3": run_c3, "c4": run_c4, "c5": run_c5}[args.workload](args)))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(row-block sharding across GPUs is not built)", file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    ctx = sa.Context(0)
    m = icosphere(1.0, wl["subdiv"])
    n = m.num_elements
    k = wl["k"]
    ph = sa.PhysicsParams.from_wavenumber(k)
    beta = ph.burton_miller_beta()
    t0 = time.perf_counter()
    tm = sa.TbemMesh(ctx, m)
    ws = sa.TbemAssemblyWorkspace(ctx, tm)
    ctx.synchronize()
    setup_ms = (time.perf_counter() - t0) * 1e3
    st = ws.stats
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    x = sa.DeviceVector(ctx, n)
    inc = sa.IncidentField.plane_wave_z()
    cfg = sa.GmresConfig(max_iterations=100, restart=wl["restart"], tolerance=1e-8)
    rhs_ptr = system.rhs_pointer()

    def step():
        ws.assemble(ph, beta, system)
        inc.add_to_system(system, tm, ph, beta, sign=1.0)
        return sa.gmres_device(op, rhs_ptr, x, cfg)

    for _ in range(args.warmup):
        sol = step()
    # Timed region: kernel events on, taken from the context's pool (the
    # mixed-precision leg below is timed the same way).
    ctx.set_profiling(True)
    ctx.reset_stats()
    ctx.synchronize()
    t0 = time.perf_counter()
    sols = [step() for _ in range(args.steps)]
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    stats = {name: ctx.kernel_stats(name) for name in sa._ffi.KERNEL_CLASSES}
    ctx.set_profiling(False)
    # the same steps with no events at all, to show what the events cost
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    elapsed_no_events = time.perf_counter() - t0
    sol = sols[-1]
    if not all(s.converged and s.residual <= 1e-8 for s in sols):
        raise SystemExit(f"bench: a timed solve did not converge: {[(s.status, s.residual) for s in sols]}")
    ms_per_step = elapsed / args.steps * 1e3
    value = args.steps / elapsed

    far_launches, far_ms = stats["assemble_far"]
    mv_launches, mv_ms = stats["matvec"]
    far_pairs = n * n - n - st["near_pairs"]
    far_flops = far_pairs * (13 * FLOPS_PER_EVAL + FLOPS_PER_COEFF)
    far_avg_s = far_ms / max(far_launches, 1) / 1e3
    mv_bytes = 16 * n * n + 32 * n
    mv_avg_s = mv_ms / max(mv_launches, 1) / 1e3
    kernels = {
        "assemble_far": {"launches": far_launches, "avg_ms": far_avg_s * 1e3,
                         "tflops": far_flops / far_avg_s / 1e12, "flops_per_launch": far_flops},
        "matvec": {"launches": mv_launches, "avg_ms": mv_avg_s * 1e3,
                   "gbs": mv_bytes / mv_avg_s / 1e9, "bytes_per_launch": mv_bytes},
        "assemble_near_ms": stats["assemble_near"][1] / max(stats["assemble_near"][0], 1),
        "assemble_self_ms": stats["assemble_self"][1] / max(stats["assemble_self"][0], 1),
        "krylov_vector_ms_total": stats["krylov_vector"][1] / args.steps,
        "rhs_ms": stats["rhs"][1] / max(stats["rhs"][0], 1),
    }
    if mv_ms >= far_ms:
        achieved = mv_bytes / mv_avg_s / 1e9
        roofline = {"kernel": "zgemv_n_kernel", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": MEASURED_TRAFFIC.get(("zgemv_n_kernel", n))}
    else:
        achieved = far_flops / far_avg_s / 1e12
        roofline = {"kernel": "far_kernel", "bound": "fp64-valu", "achieved": achieved,
                    "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                    "traffic": MEASURED_TRAFFIC.get(("far_kernel", n))}

    chosen = None
    if not args.no_chosen_form:
        x64 = x.download()
        chosen, x_ir = run_ir_leg(ctx, ws, tm, ph, beta, n, args.steps, "fp16", 1e-3)
        chosen["error_vs_fp64"] = float(np.linalg.norm(x_ir - x64) / np.linalg.norm(x64))
        chosen["vs_plain_fp64"] = chosen["value"] / value

    # matrix-free exact operator on the same workspace (apply_tbem_stabilized_exact)
    exact_op = sa.TbemExactOperator(ws, ph, beta)
    xr = np.random.default_rng(0).standard_normal(n) + 1j * np.random.default_rng(1).standard_normal(n)
    y_exact = exact_op.apply(xr)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        exact_op.apply(xr)
    ctx.synchronize()
    ex_ms = (time.perf_counter() - t0) / 3 * 1e3
    y_dense = op.apply(xr)
    exact_info = {"apply_ms": ex_ms, "tflops_far_algorithmic": far_flops / (ex_ms * 1e-3) / 1e12,
                  "rel_diff_vs_dense": float(np.linalg.norm(y_exact - y_dense) / np.linalg.norm(y_dense)),
                  "device_bytes": int(16 * (2 * n + st["near_pairs"]))}
    exact_op.close()

    # wavenumber derivative of the same system (build_tbem_system_stabilized_wavenumber_derivative)
    dsys = sa.TbemSystem(ctx, n)
    ws.assemble_wavenumber_derivative(ph, system=dsys)
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    for _ in range(2):
        ws.assemble_wavenumber_derivative(ph, system=dsys)
    ctx.synchronize()
    d_wall = (time.perf_counter() - t0) / 2
    ctx.set_profiling(False)
    d_stats = {name: ctx.kernel_stats(name) for name in ("assemble_far", "assemble_near", "assemble_self")}
    deriv_info = {"ms": d_wall * 1e3,
                  "far_kernel_avg_ms": d_stats["assemble_far"][1] / max(d_stats["assemble_far"][0], 1),
                  "near_kernel_avg_ms": d_stats["assemble_near"][1] / max(d_stats["assemble_near"][0], 1),
                  "self_kernel_avg_ms": d_stats["assemble_self"][1] / max(d_stats["assemble_self"][0], 1),
                  "value_far_kernel_avg_ms": far_avg_s * 1e3}
    del dsys

    # exterior field on a 64 x 128 sphere grid at r = 2 (generate_sphere_eval_points, pressure.rs:2154)
    th = np.pi * (np.arange(64) + 0.5) / 64
    phi = 2 * np.pi * np.arange(128) / 128
    T, P = np.meshgrid(th, phi, indexing="ij")
    pts = 2.0 * np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    x64 = x.download()
    sa.compute_total_field(pts, tm, x64, inc, ph)
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        fv = sa.compute_total_field(pts, tm, x64, inc, ph)
    ctx.synchronize()
    field_wall = (time.perf_counter() - t0) / reps
    ctx.set_profiling(False)
    f_launches, f_ms = ctx.kernel_stats("field")
    f_avg_s = f_ms / max(f_launches, 1) / 1e3
    f_flops = pts.shape[0] * n * 7 * FIELD_FLOPS_PER_QP
    field_info = {"points": int(pts.shape[0]), "panels": n, "points_per_s": pts.shape[0] / field_wall,
                  "call_ms": field_wall * 1e3, "field_kernel_avg_ms": f_avg_s * 1e3,
                  "tflops_algorithmic": f_flops / f_avg_s / 1e12,
                  "frac_fp64_peak": f_flops / f_avg_s / 1e12 / FP64_PEAK_TFLOPS,
                  "flops_per_launch": f_flops, "max_abs_p_total": float(np.max(np.abs(fv.p_total)))}

    others = None
    summary = {"c2": {"plain_fp64_rhs_per_s": value, "plain_fp64_rhs_per_s_no_events": args.steps / elapsed_no_events}}
    if chosen:
        summary["c2"].update({"chosen_form": "FP16 FGMRES-IR", "chosen_rhs_per_s": chosen["value"],
                              "chosen_vs_plain": chosen["vs_plain_fp64"],
                              "chosen_error_vs_fp64": chosen["error_vs_fp64"]})
    if not args.no_other_configs and args.workload == "c2":
        c3 = run_c3(args, ctx)
        print("bench: c3 done", file=sys.stderr, flush=True)
        c4 = run_c4(args, ctx)
        print("bench: c4 done", file=sys.stderr, flush=True)
        c5 = run_c5(args, ctx)
        print("bench: c5 done", file=sys.stderr, flush=True)
        r3, b4, r5 = c3["results"], c4["batched"], c5["results"]
        others = {"c3": c3, "c4": c4, "c5": c5}
        summary["c3"] = {"plain": "block GMRES FP64 (shared basis)",
                         "plain_rhs_per_s_excl_assembly": r3["plain_fp64"]["rhs_per_s_excl_assembly"],
                         "plain_rhs_per_s_incl_assembly": r3["plain_fp64"]["rhs_per_s_incl_assembly"],
                         "chosen": "FP16 FGMRES-IR", "chosen_rhs_per_s_excl_assembly": r3["fp16_ir"]["rhs_per_s_excl_assembly"],
                         "chosen_vs_plain_excl_assembly": r3["fp16_ir"]["vs_plain_excl_assembly"],
                         "chosen_vs_plain_incl_assembly": r3["fp16_ir"]["vs_plain_incl_assembly"],
                         "chosen_error_vs_fp64_max": r3["fp16_ir"]["error_vs_fp64_per_rhs_max"],
                         "bf16_vs_plain_excl_assembly": r3["bf16_ir"]["vs_plain_excl_assembly"],
                         "bf16_error_vs_fp64_max": r3["bf16_ir"]["error_vs_fp64_per_rhs_max"]}
        summary["c4"] = {"plain": "per-instance FP64, block GMRES", "plain_rhs_per_s": c4["plain_fp64_sequential"]["rhs_per_s"],
                         "plain_p50_ms": c4["plain_fp64_sequential"]["p50_ms"],
                         "chosen": "batched FP16 IR", "chosen_rhs_per_s": b4["batched_fp16_ir"]["rhs_per_s"],
                         "chosen_p50_ms": b4["batched_fp16_ir"]["p50_ms"], "chosen_vs_plain": b4["batched_fp16_ir"]["vs_plain"],
                         "chosen_error_vs_plain_max": b4["batched_fp16_ir"]["error_vs_plain_per_rhs_max"]}
        ch = "chosen_batched_fp16_ir_learned_then_neighbours"
        summary["c5"] = {"plain": "zero-guess FP64 GMRES", "plain_rhs_per_s": r5["zero"]["rhs_per_s"],
                         "previous_frequency_vs_plain": r5["previous"]["vs_zero"],
                         "learned_vs_plain": r5["learned"]["vs_zero"],
                         "gcro_dr_vs_plain": r5["gcro_dr"]["vs_zero"],
                         "gcro_dr_previous_vs_plain": r5["gcro_dr_previous"]["vs_zero"],
                         "chosen": "batched FP16 IR over interleaved frequency groups (frequency-family "
                                   "assembly); first group from learned guesses, later groups from "
                                   "minimal-residual combinations of the solved neighbours",
                         "chosen_rhs_per_s": r5[ch]["rhs_per_s"], "chosen_vs_plain": r5[ch]["vs_zero"],
                         "chosen_error_vs_fp64_max": r5[ch]["error_vs_fp64_max_sampled"],
                         "batched_vs_plain": {m: r5[m]["vs_zero"] for m in r5 if m.startswith(("batched", "chosen"))}}

    # measured ceilings beside the published peaks: a streaming read of a
    # matrix-sized buffer and independent FP64 FMAs on every CU
    gbs, tfl = ctypes.c_double(), ctypes.c_double()
    sa._check(sa._lib().sbh_debug_hbm_read_rate(ctx.handle, 16 * n * n, 5, ctypes.byref(gbs)))
    sa._check(sa._lib().sbh_debug_fp64_fma_rate(ctx.handle, ctypes.byref(tfl)))
    ceilings = {"hbm_stream_read_gbs": gbs.value, "fp64_fma_tflops": tfl.value,
                "matvec_frac_of_stream_read": kernels["matvec"]["gbs"] / gbs.value,
                "far_kernel_frac_of_fma_rate": kernels["assemble_far"]["tflops"] / tfl.value}

    cpu = None
    if not args.no_cpu_baseline:
        threads = host_info()["threads_used"]
        cpu = cpu_baseline(m, k, beta, sol.spmv_count, args.cpu_rows, threads)

    line = {
        "metric": "solved RHS/s at rel. residual 1e-8, chosen form vs plain FP64 on GPU; % roofline",
        "value": value,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (generated icosphere mesh, analytic plane wave)",
        "config": {"workload": wl["name"], "panels": n, "k": k, "form": "plain FP64",
                   "parallelism": "single GPU"},
        "solve": {"iterations": sol.iterations, "restarts": sol.restarts, "spmv_count": sol.spmv_count,
                  "residual": sol.residual},
        "setup_ms": setup_ms,
        "routes": st,
        "kernels": kernels,
        "roofline": roofline,
        "measured_ceilings": ceilings,
        "summary_chosen_vs_plain": summary,
        "plain_events_off_ms_per_step": elapsed_no_events / args.steps * 1e3,
        "chosen_form": chosen,
        "exact_operator": exact_info,
        "field_evaluation": field_info,
        "wavenumber_derivative_assembly": deriv_info,
        "other_configs": others,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
