"""Batched frequency sweeps on one rigid mesh.

A sweep over equally spaced wavenumbers k_0 < k_1 < ... is cut into
``stride`` interleaved groups (sweep indices g, g + stride, g + 2 stride, ...)
that are solved one batch at a time, coarse to fine: group 0 first, then the
groups that bisect the gaps between solved groups (stride 4: 0, 2, 1, 3).
Each group is one batch (sbh_batch_*) whose instances share the mesh; its
wavenumbers step by a constant, so the batch assembly takes the
frequency-family path (one pass over the pair geometry per 16 wavenumbers),
and it is solved by lock-step FP16/BF16 refinement (or FP64 GMRES) to the FP64
residual bar.  Initial guesses come from ``warmstart.batch_guesses``: the
first group from the network (or zero), later groups from the minimal-residual
combination of the solutions at the nearest solved sweep points on either
side, taken from the two groups that bracket the group in the order (the
reference's sweep carries the previous solution only,
frequency_sweep.rs:718-730).  Solved points stay in a device store; guesses
are formed on the device (sbh_batch_learned_guesses, sbh_batch_guess_gram /
sbh_batch_guess_combine, sbh_batch_set_guesses_device) and only the
solutions of the whole sweep come back to the host, at the end.

Tried and not kept: two contexts on the device, one host thread each, so that
one group's assembly overlaps another's solve.  The persistent assembly
launch holds every CU while it runs, and the solve's many short launches
queued behind it: a 32-instance solve took 300 ms beside a concurrent
assembly against 142 ms alone, more than the two run one after the other.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import Batch, Context, DeviceVector, TbemMesh
from . import warmstart as wsm

__all__ = ["group_order", "group_neighbours", "SweepResult", "SweepBatch", "batched_sweep"]


def group_order(stride: int) -> List[int]:
    """Coarse-to-fine order of the interleaved groups: 0, then bisection."""
    if stride < 1 or stride & (stride - 1):
        raise ValueError("stride must be a power of two")
    out, step = [0], stride
    while step > 1:
        step //= 2
        out += list(range(step, stride, 2 * step))
    return out


def group_neighbours(stride: int):
    """For each group offset g, the offsets (lower, upper) of the groups solved
    before it whose sweep points bracket g's points (upper may be ``stride``,
    i.e. group 0 shifted by one period); None for the first group."""
    order = group_order(stride)
    nb = {order[0]: None}
    for pos, g in enumerate(order[1:], start=1):
        earlier = order[:pos]
        lo = max(e for e in earlier if e < g)
        his = [e for e in earlier if e > g]
        hi = min(his) if his else stride
        nb[g] = (lo, hi)
    return nb


@dataclass
class SweepResult:
    solutions: np.ndarray                 # len(ks) x n
    iterations: np.ndarray                # Arnoldi steps per sweep point
    residuals: np.ndarray                 # recomputed FP64 relative residuals
    guess_residuals: dict = field(default_factory=dict)  # sweep index -> guess's relative residual
    seconds: float = 0.0
    breakdown_s: dict = field(default_factory=dict)


class SweepBatch:
    """The batch of one group (``group`` instances on ``mesh``), kept between
    sweeps, with the device store of solved sweep points and a buffer for the
    network's guesses."""

    def __init__(self, ctx: Context, mesh, group: int):
        self.ctx = ctx
        self.group = int(group)
        self.n = mesh.num_elements
        self.tm = TbemMesh(ctx, mesh)
        self.c, self.nn, _ = self.tm.geometry()
        self.batch = Batch(ctx, [self.tm] * self.group)
        self.store = None
        self.learned = None

    def ensure_store(self, points: int) -> DeviceVector:
        if self.store is None or self.store.length < points * self.n:
            if self.store is not None:
                self.store.close()
            self.store = DeviceVector(self.ctx, points * self.n)
        if self.learned is None:
            self.learned = DeviceVector(self.ctx, self.group * self.n)
        return self.store

    def pointer(self, j: int) -> int:
        """Device address of sweep point j's solution in the store."""
        return self.store.ptr + 16 * self.n * int(j)

    def close(self):
        for v in (self.store, self.learned):
            if v is not None:
                v.close()
        self.batch.close()
        self.tm.close()


def batched_sweep(ctx: Context, mesh, ks, direction, *, stride: int = 4, net: Optional["wsm.WarmStartMLP"] = None,
                  first_guess: Optional[str] = "learned", later_guess: Optional[str] = "neighbours",
                  fmt: str = "fp16", restart: int = 30, max_cycles: int = 30, tolerance: float = 1e-8,
                  inner_tolerance: float = 1e-3, products: str = "fp16", per_side: int = 1,
                  sweep_batch: Optional[SweepBatch] = None, use_graph: bool = True) -> SweepResult:
    """Solve the rigid-mesh sweep ``ks`` (equally spaced, len divisible by
    ``stride``) for one plane wave along ``direction``.  first_guess: None or
    "learned"; later_guess: None, "previous", "learned", "neighbours" or
    "neighbours+learned" (see ``warmstart.batch_guesses``; ``products`` is the
    format of the batched products the combinations use, ``per_side`` the
    solved points per side).  ``sweep_batch`` lets a caller keep the batch
    between sweeps (its setup, the routes, is then outside the timing);
    ``use_graph`` False launches every solver kernel directly instead of
    replaying the captured cycle."""
    ks = np.asarray(ks, dtype=np.float64)
    if len(ks) % stride:
        raise ValueError("the number of wavenumbers must be a multiple of stride")
    uses_net = any(kind is not None and "learned" in kind for kind in (first_guess, later_guess))
    if uses_net and net is None:
        raise ValueError("learned guesses need a network")
    group = len(ks) // stride
    own = sweep_batch is None
    sb = SweepBatch(ctx, mesh, group) if own else sweep_batch
    if sb.group != group:
        raise ValueError("the sweep batch holds a different number of instances")
    n = mesh.num_elements
    d = np.asarray(direction, dtype=np.float64).reshape(3)
    neighbours = group_neighbours(stride)
    iters = np.zeros(len(ks), dtype=np.int64)
    res = np.zeros(len(ks))
    guess_res = {}
    parts = {"assemble_convert": 0.0, "guesses": 0.0, "solve": 0.0}
    batch = sb.batch
    sb.ensure_store(len(ks))
    ctx.synchronize()
    t0 = time.perf_counter()
    try:
        for g in group_order(stride):
            idx = list(range(g, len(ks), stride))
            kg = ks[idx]
            ta = time.perf_counter()
            batch.set_problem([float(k) for k in kg], [d.reshape(1, 3)] * group)
            batch.assemble()
            if fmt != "fp64":
                batch.convert(fmt)
            ctx.synchronize()
            tb = time.perf_counter()
            kind = first_guess if neighbours[g] is None else later_guess
            if kind is not None:
                # only the bracketing groups count
                allowed = set() if neighbours[g] is None else {neighbours[g][0], neighbours[g][1] % stride}
                solved = {j for j in range(len(ks)) if j % stride in allowed}
                rels = wsm.batch_guesses_device(batch, kind, idx, solved, sb.pointer, net, products=products,
                                                per_side=per_side, learned_buffer=sb.learned)
                for j, r in zip(idx, rels):
                    guess_res[j] = r
            tc = time.perf_counter()
            sts = batch.solve(fmt if fmt == "fp64" else fmt + "_ir", restart=restart, max_cycles=max_cycles,
                              tolerance=tolerance, inner_tolerance=inner_tolerance, use_graph=use_graph,
                              use_guess=kind is not None)
            if not all(q.converged for q in sts):
                raise RuntimeError(f"sweep group {g} left a frequency unconverged")
            batch.store_solutions([sb.pointer(j) for j in idx])
            for i, j in enumerate(idx):
                iters[j] = sts[i].iterations
                res[j] = sts[i].residual
            td = time.perf_counter()
            parts["assemble_convert"] += tb - ta
            parts["guesses"] += tc - tb
            parts["solve"] += td - tc
        X = sb.store.download(0, len(ks) * n).reshape(len(ks), n)
        ctx.synchronize()
        el = time.perf_counter() - t0
    finally:
        if own:
            sb.close()
    return SweepResult(X, iters, res, guess_res, el, parts)
