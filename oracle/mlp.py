"""CPU oracle of the warm-start MLP (test infrastructure only).

A numpy restatement of the network the HIP kernels train: features of a
panel sample, four linear layers with SiLU between them, the loss
sum |y - t|^2 / (2M) and its gradients.  With ``bf16=True`` it rounds to
BF16 (round to nearest even) exactly where the kernels store BF16 values:
the features, the weights, the hidden activations and every dZ; sums are in
float64, so it differs from the kernels only by FP32 accumulation order and
by the rare BF16 rounding flips that follow from it.  With ``bf16=False`` it
is the plain float64 network, whose gradients the tests check against
central finite differences.

Parameter layout (one vector): per layer the row-major weights (out x in)
followed by the bias (out); layer sizes 64 -> W -> W -> W -> 2.
"""

from __future__ import annotations

import numpy as np

FEAT = 64


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    b = f.view(np.uint32).astype(np.uint64)
    rounded = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def features(raw):
    """raw: (M, 10) = c[3], n[3], d[3], k.  Returns (M, 64) float64."""
    raw = np.asarray(raw, dtype=np.float32)
    c, n, d, k = raw[:, 0:3], raw[:, 3:6], raw[:, 6:9], raw[:, 9]
    kk = (k * np.float32(1.0 / 12.0)).astype(np.float32)
    dc = (d * c).sum(axis=1, dtype=np.float32)
    dn = (d * n).sum(axis=1, dtype=np.float32)
    cols = [c[:, 0], c[:, 1], c[:, 2], n[:, 0], n[:, 1], n[:, 2], d[:, 0], d[:, 1], d[:, 2], kk, dn, dc]
    for q in (dc, dn, kk, (kk * dc).astype(np.float32)):
        w = np.float32(3.14159265358979)
        for _ in range(6):
            cols.append(np.sin(w * q, dtype=np.float32))
            cols.append(np.cos(w * q, dtype=np.float32))
            w = np.float32(w * 2)
    out = np.zeros((raw.shape[0], FEAT), dtype=np.float64)
    out[:, : len(cols)] = np.stack(cols, axis=1)
    return out


def layer_shapes(width):
    return [(width, FEAT), (width, width), (width, width), (2, width)]


def unpack(params, width):
    p = np.asarray(params, dtype=np.float64)
    out, off = [], 0
    for rows, cols in layer_shapes(width):
        w = p[off: off + rows * cols].reshape(rows, cols)
        off += rows * cols
        b = p[off: off + rows]
        off += rows
        out.append((w, b))
    assert off == p.size
    return out


def num_params(width):
    return sum(r * c + r for r, c in layer_shapes(width))


def silu(x):
    return x / (1.0 + np.exp(-x))


def silu_grad(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def loss_and_gradients(params, width, raw, targets, bf16_storage=True):
    """Returns (loss, grads, predictions) for the samples in raw/targets."""
    rnd = bf16 if bf16_storage else (lambda v: np.asarray(v, dtype=np.float64))
    layers = unpack(params, width)
    X = rnd(features(raw))
    M = X.shape[0]
    acts, pre = [X], []
    a = X
    for li, (w, b) in enumerate(layers[:-1]):
        h = a @ rnd(w).T + b
        if bf16_storage:
            h = h.astype(np.float32).astype(np.float64)
        pre.append(h)
        a = rnd(silu(h))
        acts.append(a)
    w4, b4 = layers[-1]
    y = acts[-1] @ rnd(w4).T + b4
    t = np.asarray(targets, dtype=np.float64)
    e = y - t
    loss = 0.5 * float((e * e).sum()) / M
    dz = rnd(e / M)
    grads = [None] * len(layers)
    for li in range(len(layers) - 1, -1, -1):
        w, _ = layers[li]
        grads[li] = (dz.T @ acts[li], dz.sum(axis=0))
        if li == 0:
            break
        da = dz @ rnd(w)
        dz = rnd(da * silu_grad(pre[li - 1]))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat, y


def predict(params, width, raw, bf16_storage=True):
    t = np.zeros((np.asarray(raw).shape[0], 2))
    return loss_and_gradients(params, width, raw, t, bf16_storage)[2]
