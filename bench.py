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
dominant kernel (timed with HIP events on the library's stream), and the CPU
baseline (the oracle's C++ restatement of the reference algorithm timed on a
sample of rows on this host's cores).
"""

from __future__ import annotations

import argparse
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

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# this workload (copy-calibrated; profiles/zgemv_n_kernel_and_far_kernel_c2_icosphere20480_k10.md).
MEASURED_TRAFFIC = {
    ("zgemv_n_kernel", 20480): 6.7134e9 + 3.2971e5,
    ("far_kernel", 20480): 1.8935e7 + 6.7109e9,
}

WORKLOADS = {
    "c2": {"subdiv": 5, "k": 10.0, "restart": 100,
           "name": "rigid sphere icosphere 20480 panels, k=10, beta=i/k, 1 plane wave +z, GMRES(100) rtol 1e-8, complex128"},
    "c3": {"name": "many right-hand sides (separate report)"},
    "c1": {"subdiv": 3, "k": 2.0, "restart": 50,
           "name": "rigid sphere icosphere 1280 panels, k=2, beta=i/k, 1 plane wave +z, GMRES(50) rtol 1e-8, complex128"},
}


def cpu_baseline(m, k, beta, spmv_count, rows, threads):
    """Oracle (C++ restatement of the reference CPU path) on `rows` rows; extrapolated to N."""
    import oracle

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
    return {
        "value": 1.0 / t_total,
        "unit": "RHS/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle assembly of rows [0,{rows}) of {n} and {reps} matvecs of that row block, "
                   f"extrapolated x{scale:.1f}; solve = {spmv_count} matvecs (the GPU run's count); "
                   f"assembly {t_asm * scale:.1f} s + solve {t_solve:.1f} s per RHS"),
    }


def run_ir_leg(ctx, ws, tm, ph, beta, n, steps, fmt, inner_tol):
    """Mixed-precision form of the same step: assemble FP64, convert to the
    row-scaled complex FP16/BF16 copy, FGMRES-IR to an FP64-recomputed
    residual of 1e-8."""
    system = sa.TbemSystem(ctx, n)
    op = sa.DenseMatrixOperator(system)
    inc = sa.IncidentField.plane_wave_z()
    cfg = sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8, max_inner_iterations=50,
                           inner_tolerance=inner_tol)
    ws.assemble(ph, beta, system)
    lp = sa.LowpMatrix(system, fmt)

    def step():
        ws.assemble(ph, beta, system)
        lp.update()
        b = inc.compute_rhs_with_beta(tm, ph, beta)
        return sa.gmres_ir(op, lp, b, cfg)

    step()
    ctx.synchronize()
    t0 = time.perf_counter()
    outs = [step() for _ in range(steps)]
    ctx.synchronize()
    el = time.perf_counter() - t0
    x, st = outs[-1]
    if not all(o[1][0].converged and o[1][0].final_outer_residual <= 1e-8 for o in outs):
        raise SystemExit(f"bench: mixed-precision solve did not converge: {st[0]}")
    return {"form": f"{fmt} storage + FP32 accumulate, FGMRES-IR with FP64 residual/Arnoldi", "value": steps / el,
            "ms_per_step": el / steps * 1e3, "outer_iterations": st[0].outer_iterations,
            "inner_iterations": st[0].total_inner_iterations, "residual": st[0].final_outer_residual,
            "fp64_products": st[0].fp64_products, "lowp_products": st[0].lowp_products}, x[:, 0]


def run_c3(args):
    """Many right-hand sides: 40960-panel Fibonacci sphere, k = 14, 64 plane waves."""
    ctx = sa.Context(0)
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
    t0 = time.perf_counter()
    ws.assemble(ph, beta, system)
    ctx.synchronize()
    t_asm = time.perf_counter() - t0
    B = inc.compute_rhs_with_beta(tm, ph, beta)
    res = {}
    ctx.reset_stats()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    X64, sols = sa.gmres_block(op, B, sa.GmresConfig(max_iterations=100, restart=100, tolerance=1e-8))
    t_plain = time.perf_counter() - t0
    blk = ctx.kernel_stats("block")
    res["plain_fp64"] = {"solve_s": t_plain, "rhs_per_s_excl_assembly": 64 / t_plain,
                         "rhs_per_s_incl_assembly": 64 / (t_plain + t_asm),
                         "max_iterations": max(s.iterations for s in sols),
                         "max_residual": max(s.residual for s in sols),
                         "fp64_block_product_avg_ms": blk[1] / max(blk[0], 1)}
    for fmt, itol in (("fp16", 1e-3), ("bf16", 1e-2)):
        ctx.reset_stats()
        t0 = time.perf_counter()
        lp = sa.LowpMatrix(system, fmt)
        ctx.synchronize()
        t_conv = time.perf_counter() - t0
        t0 = time.perf_counter()
        X, st = sa.gmres_ir(op, lp, B, sa.GmresIrConfig(max_outer_iterations=30, outer_tolerance=1e-8,
                                                         max_inner_iterations=50, inner_tolerance=itol))
        t_ir = time.perf_counter() - t0
        lpk = ctx.kernel_stats("lowp")
        blk = ctx.kernel_stats("block")
        err = [float(np.linalg.norm(X[:, c] - X64[:, c]) / np.linalg.norm(X64[:, c])) for c in range(64)]
        lp_avg = lpk[1] / max(lpk[0], 1) / 1e3
        lp_bytes = 4 * n * n
        lp_flops = 8 * n * n * 64
        res[fmt + "_ir"] = {"solve_s": t_ir, "convert_s": t_conv,
                            "rhs_per_s_excl_assembly": 64 / (t_ir + t_conv),
                            "rhs_per_s_incl_assembly": 64 / (t_ir + t_conv + t_asm),
                            "max_outer": max(s.outer_iterations for s in st),
                            "max_inner": max(s.total_inner_iterations for s in st),
                            "max_residual": max(s.final_outer_residual for s in st),
                            "max_error_vs_fp64": max(err), "converged": all(s.converged for s in st),
                            "lowp_gemm_avg_ms": lp_avg * 1e3, "lowp_gemm_gbs": lp_bytes / lp_avg / 1e9,
                            "lowp_gemm_tflops": lp_flops / lp_avg / 1e12,
                            "fp64_block_product_avg_ms": blk[1] / max(blk[0], 1)}
        lp.close()
    ctx.set_profiling(False)
    line = {"metric": "solved RHS/s at rel. residual 1e-8 (multi-RHS)", "workload": "C3: Fibonacci sphere 40960 panels, k=14, 64 plane waves",
            "panels": n, "assembly_s": t_asm, "results": res}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="c2")
    ap.add_argument("--cpu-rows", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-chosen-form", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run 3 plain device copies of BYTES bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        ctx = sa.Context(0)
        sa._check(sa._lib().sbh_debug_device_copy(ctx.handle, args.calibrate_copy, 3))
        print(json.dumps({"copy_bytes": args.calibrate_copy, "launches": 3}))
        return
    if args.workload == "c3":
        run_c3(args)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU row sharding is described in DESIGN.md but not built)", file=sys.stderr)
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
    _, _, rhs_ptr = system.device_pointers()

    def step():
        ws.assemble(ph, beta, system)
        inc.add_to_system(system, tm, ph, beta, sign=1.0)
        return sa.gmres_device(op, rhs_ptr, x, cfg)

    for _ in range(args.warmup):
        sol = step()
    ctx.synchronize()
    ctx.reset_stats()
    ctx.set_profiling(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    sols = [step() for _ in range(args.steps)]
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.set_profiling(False)
    sol = sols[-1]
    if not all(s.converged and s.residual <= 1e-8 for s in sols):
        raise SystemExit(f"bench: a timed solve did not converge: {[(s.status, s.residual) for s in sols]}")
    ms_per_step = elapsed / args.steps * 1e3
    value = args.steps / elapsed
    stats = {name: ctx.kernel_stats(name) for name in sa._ffi.KERNEL_CLASSES}

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

    cpu = None
    if not args.no_cpu_baseline:
        threads = max(1, min(os.cpu_count() or 1, 16))
        cpu = cpu_baseline(m, k, beta, sol.spmv_count, args.cpu_rows, threads)

    line = {
        "metric": "solved RHS/s at rel. residual 1e-8 (plain FP64 on GPU)",
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
        "chosen_form": chosen,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
