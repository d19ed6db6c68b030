"""Learned warm start for frequency sweeps.

The network (HIP kernels in csrc/mlp_kernels.hip) maps one panel sample
{centroid, outward normal, incident direction, k} to the incident-phase-free
surface density u = x exp(-i k d.c); the GMRES initial guess is
x0 = u exp(i k d.c).  The reference's own warm start is the previous
frequency's solution (frequency_sweep.rs:718-730), passed to gmres_with_guess
(gmres.rs:564-626); the zero-guess solve stays the fallback: a guess whose
residual is not below ||b|| is replaced by zero, and every solve is judged by
the recomputed FP64 residual.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _ffi
from . import _check, _cptr, _lib, Context, DenseMatrixOperator, GmresConfig, gmres_with_guess, gmres

RAW = 10


def panel_inputs(centers, normals, k: float, direction) -> np.ndarray:
    """(N, 10) float32 raw inputs of one solve (one incident direction)."""
    n = centers.shape[0]
    out = np.empty((n, RAW), dtype=np.float32)
    out[:, 0:3] = centers
    out[:, 3:6] = normals
    out[:, 6:9] = np.asarray(direction, dtype=np.float64).reshape(1, 3)
    out[:, 9] = k
    return out


def incident_phase(centers, k: float, direction) -> np.ndarray:
    return np.exp(1j * k * (centers @ np.asarray(direction, dtype=np.float64).reshape(3)))


def incident_phases(centers, ks, direction) -> np.ndarray:
    """exp(i k d.c) for every k of ks (len(ks) x N).  For equally spaced ks the
    rows follow from the first by repeated products with exp(i dk d.c)
    (relative error about one ulp per row), else one complex exp each."""
    ks = np.asarray(ks, dtype=np.float64)
    dc = centers @ np.asarray(direction, dtype=np.float64).reshape(3)
    out = np.empty((len(ks), dc.shape[0]), dtype=np.complex128)
    if len(ks) > 2 and np.allclose(np.diff(ks), ks[1] - ks[0], rtol=1e-12, atol=0.0):
        ph = np.exp(1j * ks[0] * dc)
        step = np.exp(1j * (ks[1] - ks[0]) * dc)
        for f in range(len(ks)):
            out[f] = ph
            ph = ph * step
    else:
        out[:] = np.exp(1j * ks[:, None] * dc[None, :])
    return out


def density_to_target(x, centers, k: float, direction) -> np.ndarray:
    """u = x exp(-i k d.c) as (N, 2) float32."""
    u = np.asarray(x) * np.conj(incident_phase(centers, k, direction))
    return np.stack([u.real, u.imag], axis=1).astype(np.float32)


def target_to_density(y, centers, k: float, direction) -> np.ndarray:
    u = np.asarray(y[:, 0], dtype=np.float64) + 1j * np.asarray(y[:, 1], dtype=np.float64)
    return u * incident_phase(centers, k, direction)


class WarmStartMLP:
    """4-layer MLP (64 -> width -> width -> width -> 2), BF16 weights, FP32 accumulation."""

    def __init__(self, ctx: Context, width: int = 512, seed: int = 1):
        self.ctx = ctx
        self.width = int(width)
        h = ctypes.c_void_p()
        _check(_lib().sbh_mlp_create(ctx.handle, self.width, int(seed), ctypes.byref(h)))
        self._h = h
        self.num_params = int(_lib().sbh_mlp_num_params(h))

    def params(self) -> np.ndarray:
        out = np.empty(self.num_params, dtype=np.float32)
        _check(_lib().sbh_mlp_get_params(self._h, _cptr(out)))
        return out

    def set_params(self, p) -> None:
        v = np.ascontiguousarray(p, dtype=np.float32).reshape(self.num_params)
        _check(_lib().sbh_mlp_set_params(self._h, _cptr(v)))

    def loss_and_gradients(self, inputs, targets):
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(self.num_params, dtype=np.float32)
        loss = ctypes.c_double()
        _check(_lib().sbh_mlp_loss_and_gradients(self._h, _cptr(x), _cptr(t), x.shape[0], ctypes.byref(loss),
                                                 _cptr(g)))
        return loss.value, g

    def set_dataset(self, inputs, targets) -> None:
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        t = np.ascontiguousarray(targets, dtype=np.float32)
        _check(_lib().sbh_mlp_set_dataset(self._h, _cptr(x), _cptr(t), x.shape[0]))

    def train(self, steps: int, batch: int = 16384, lr_max: float = 2e-3, lr_min: float = 5e-5,
              seed: int = 1) -> np.ndarray:
        losses = np.zeros(int(steps))
        _check(_lib().sbh_mlp_train(self._h, int(steps), int(batch), float(lr_max), float(lr_min), int(seed),
                                    _cptr(losses)))
        return losses

    def predict(self, inputs) -> np.ndarray:
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        out = np.empty((x.shape[0], 2), dtype=np.float32)
        _check(_lib().sbh_mlp_predict(self._h, _cptr(x), x.shape[0], _cptr(out), 0))
        return out

    def predict_layered(self, inputs) -> np.ndarray:
        """predict() through the layer-by-layer forward pass of training (test hook)."""
        x = np.ascontiguousarray(inputs, dtype=np.float32)
        out = np.empty((x.shape[0], 2), dtype=np.float32)
        _check(_lib().sbh_debug_mlp_predict_layered(self._h, _cptr(x), x.shape[0], _cptr(out)))
        return out

    def guess(self, centers, normals, k: float, direction) -> np.ndarray:
        """GMRES initial guess x0 for one (k, direction) on a mesh."""
        y = self.predict(panel_inputs(centers, normals, k, direction))
        return target_to_density(y, centers, k, direction)

    def guesses(self, centers, normals, ks, direction) -> np.ndarray:
        """Initial guesses for several wavenumbers in one forward pass: (len(ks), N)."""
        n = centers.shape[0]
        raw = np.empty((len(ks), n, RAW), dtype=np.float32)
        raw[:, :, 0:3] = centers[None]
        raw[:, :, 3:6] = normals[None]
        raw[:, :, 6:9] = np.asarray(direction, dtype=np.float64).reshape(1, 1, 3)
        raw[:, :, 9] = np.asarray(ks, dtype=np.float64)[:, None]
        raw = raw.reshape(-1, RAW)
        y = self.predict(raw).reshape(len(ks), n, 2)
        return (y[..., 0] + 1j * y[..., 1]) * incident_phases(centers, ks, direction)

    def close(self):
        if getattr(self, "_h", None):
            _lib().sbh_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_with_guess(op: DenseMatrixOperator, b, x0: Optional[np.ndarray], config: GmresConfig,
                     with_residual: bool = False):
    """gmres_with_guess with the zero-guess fallback: a guess whose residual is not
    below ||b|| is dropped.  Returns (solution, used_guess), plus the guess's
    relative residual ||b - A x0|| / ||b|| (1.0 without a guess) when asked."""
    rel = 1.0
    sol = None
    if x0 is not None:
        r0 = np.linalg.norm(b - op.apply(x0))
        rel = float(r0 / np.linalg.norm(b))
        if np.isfinite(r0) and rel < 1.0:
            sol = gmres_with_guess(op, b, x0, config)
    used = sol is not None
    if sol is None:
        sol = gmres(op, b, config)
    return (sol, used, rel) if with_residual else (sol, used)


def minimal_residual_combination(b, candidates, products):
    """x0 = sum_j c_j x_j with c minimising ||b - sum_j c_j (A x_j)||.

    candidates and products are lists of length-n vectors (A x_j computed by
    the caller, on the device); an all-zero candidate gets coefficient zero.
    Returns (x0, ||b - A x0|| / ||b||).  A minimal-residual combination of
    earlier solutions generalises the previous-frequency guess
    (frequency_sweep.rs:718-730): with one candidate it is that guess scaled
    by its optimal factor."""
    P = np.stack([np.asarray(p, dtype=np.complex128) for p in products], axis=1)
    X = np.stack([np.asarray(x, dtype=np.complex128) for x in candidates], axis=1)
    b = np.asarray(b, dtype=np.complex128)
    coef, *_ = np.linalg.lstsq(P, b, rcond=None)
    r = b - P @ coef
    return X @ coef, float(np.linalg.norm(r) / np.linalg.norm(b))


def batch_guesses_device(batch, kind: str, indices, solved, store_pointer, net: Optional["WarmStartMLP"] = None,
                         products: str = "fp64", per_side: int = 1, learned_buffer=None):
    """batch_guesses with every vector on the device: the solved sweep points
    live in device memory (``store_pointer(j)`` is the device address of sweep
    point j's solution block; ``solved`` the solved indices), the network's
    guesses come from sbh_batch_learned_guesses, and the minimal-residual
    combinations from sbh_batch_guess_gram / sbh_batch_guess_combine; only the
    per-column m x m normal equations are solved on the host (pseudo-inverse,
    as batch_guesses).  ``learned_buffer`` is a DeviceVector of the batch's
    packed length, needed for "neighbours+learned".  Returns the
    combinations' relative residuals ||b - A x0|| / ||b|| (empty for
    "learned" and "previous")."""
    rels = []
    if kind == "learned":
        batch.learned_guesses(net)
        return rels
    if kind == "previous":
        batch.set_guesses_device([None if lo is None else store_pointer(lo)
                                  for lo, _ in (nearest_solved(solved, j) for j in indices)])
        return rels
    if kind not in ("neighbours", "neighbours+learned"):
        raise ValueError(f"unknown guess kind {kind!r}")
    near = [nearest_solved(solved, j, per_side) for j in indices]
    if per_side == 1:
        near = [([lo] if lo is not None else [], [hi] if hi is not None else []) for lo, hi in near]
    slots = []
    for r in range(per_side):
        slots.append([store_pointer(lo[r]) if r < len(lo) else None for lo, _ in near])
        slots.append([store_pointer(hi[r]) if r < len(hi) else None for _, hi in near])
    if kind == "neighbours+learned":
        if learned_buffer is None:
            raise ValueError("neighbours+learned needs a learned_buffer")
        batch.learned_guesses(net, out=learned_buffer)
        offs = np.cumsum([0] + [a * b for a, b in batch._packed_shapes()])[:-1]
        slots.append([learned_buffer.ptr + 16 * int(o) for o in offs])
    G, h, bb = batch.guess_gram(slots, products)
    coef = (np.linalg.pinv(G, rcond=1e-13, hermitian=True) @ h[:, :, None])[:, :, 0]
    # ||b - P c||^2 = ||b||^2 - 2 Re(c^H h) + c^H G c
    r2 = bb - 2.0 * np.real(np.einsum("ca,ca->c", coef.conj(), h)) + \
        np.real(np.einsum("ca,cab,cb->c", coef.conj(), G, coef))
    rel_all = np.sqrt(np.maximum(r2, 0.0) / bb)
    for c in range(coef.shape[0]):
        rel = float(rel_all[c])
        if not (np.all(np.isfinite(coef[c])) and np.isfinite(rel) and rel < 1.0):
            coef[c] = 0.0
            rel = 1.0
        rels.append(rel)
    batch.guess_combine(coef)
    return rels


def nearest_solved(solved, j: int, per_side: int = 1):
    """Indices of the nearest solved sweep points (a dict or set of indices)
    below and above j (None when absent); with per_side > 1, lists of up to
    that many on each side, nearest first."""
    below = sorted((i for i in solved if i < j), reverse=True)
    above = sorted(i for i in solved if i > j)
    if per_side == 1:
        return (below[0] if below else None), (above[0] if above else None)
    return below[:per_side], above[:per_side]


def batch_guesses(batch, kind: str, indices, solved: dict, net: Optional["WarmStartMLP"], centers, normals, ks,
                  direction, products: str = "fp64", per_side: int = 1):
    """Initial guesses for a batch whose instance i is sweep point indices[i]
    (one right-hand side each), set on the batch for solve(use_guess=True).

    kind:
      "learned"             the network's guess;
      "previous"            the solution at the nearest solved point below (zero if none);
      "neighbours"          minimal-residual combination of the solutions at the nearest solved
                            points below and above (``per_side`` of them on each side);
      "neighbours+learned"  the same with the network's guess as a third candidate.
    The products A_i x for the combinations come from one batched launch per
    candidate slot (sbh_batch_apply).  A combination whose residual is not
    below ||b|| falls back to zero.  Returns the guesses (len(indices), n) and
    the combinations' relative residuals."""
    n = centers.shape[0]
    nb = len(indices)
    zeros = np.zeros(n, dtype=np.complex128)
    learned = net.guesses(centers, normals, ks, direction) if "learned" in kind else None
    rels = []
    if kind == "learned":
        X0 = learned
    elif kind == "previous":
        X0 = np.stack([solved[lo] if lo is not None else zeros
                       for lo, _ in (nearest_solved(solved, j) for j in indices)])
    elif kind in ("neighbours", "neighbours+learned"):
        slots = []
        near = [nearest_solved(solved, j, per_side) for j in indices]
        if per_side == 1:
            near = [([lo] if lo is not None else [], [hi] if hi is not None else []) for lo, hi in near]
        for r in range(per_side):
            slots.append([solved[lo[r]] if r < len(lo) else zeros for lo, _ in near])
            slots.append([solved[hi[r]] if r < len(hi) else zeros for _, hi in near])
        if learned is not None:
            slots.append([learned[i] for i in range(nb)])
        prods = [batch.apply([x.reshape(n, 1) for x in slot], products) for slot in slots]
        # all instances at once: the m x m normal equations of every instance
        # (m = 2 per_side, + 1 with the learned guess) from batched products,
        # solved by pseudo-inverse
        Q = np.stack([np.stack([p[i][:, 0] for p in prods]) for i in range(nb)])  # nb x m x n: A_i x_c
        Xc = np.stack([np.stack([sl[i] for sl in slots]) for i in range(nb)])    # nb x m x n: x_c
        Bv = np.stack([r[:, 0] for r in batch.rhs_all()])                        # nb x n
        G = Q.conj() @ np.swapaxes(Q, 1, 2)
        h = (Q.conj() @ Bv[:, :, None])[:, :, 0]
        coef = (np.linalg.pinv(G, rcond=1e-13, hermitian=True) @ h[:, :, None])[:, :, 0]
        X0 = (np.swapaxes(Xc, 1, 2) @ coef[:, :, None])[:, :, 0]
        R = Bv - (np.swapaxes(Q, 1, 2) @ coef[:, :, None])[:, :, 0]
        rel_all = np.linalg.norm(R, axis=1) / np.linalg.norm(Bv, axis=1)
        for i in range(nb):
            rel = float(rel_all[i])
            if not (np.isfinite(rel) and rel < 1.0):
                X0[i], rel = zeros, 1.0
            rels.append(rel)
    else:
        raise ValueError(f"unknown guess kind {kind!r}")
    batch.set_guesses([x.reshape(n, 1) for x in X0])
    return X0, rels
