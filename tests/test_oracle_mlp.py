"""Pins of the MLP oracle (CPU): the backward pass against central finite
differences of the float64 network, BF16 rounding against torch's
float32 -> bfloat16 conversion (round to nearest even), and the feature map's
identities."""

import numpy as np
import pytest
import torch

from oracle import mlp


def _raw(rng, m):
    c = rng.standard_normal((m, 3))
    n = rng.standard_normal((m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    k = rng.uniform(1, 12, size=(m, 1))
    return np.concatenate([c, n, d, k], axis=1).astype(np.float32)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    width = 8
    p = rng.standard_normal(mlp.num_params(width)) * 0.3
    raw = _raw(rng, 5)
    t = rng.standard_normal((5, 2))
    loss, g, _ = mlp.loss_and_gradients(p, width, raw, t, bf16_storage=False)
    # sample parameters from every layer's weights and bias
    picks = rng.choice(p.size, size=60, replace=False)
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_and_gradients(pp, width, raw, t, False)[0] - mlp.loss_and_gradients(pm, width, raw, t, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 10000),
                        np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0], dtype=np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp.bf16(x), ref)
    assert mlp.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0  # tie to even


def test_feature_map_identities():
    rng = np.random.default_rng(5)
    raw = _raw(rng, 64)
    f = mlp.features(raw)
    assert f.shape == (64, 64)
    assert np.allclose(f[:, :9], raw[:, :9])
    assert np.allclose(f[:, 9], raw[:, 9] / 12.0, rtol=1e-6)
    d, c, n = raw[:, 6:9].astype(np.float64), raw[:, 0:3].astype(np.float64), raw[:, 3:6].astype(np.float64)
    assert np.allclose(f[:, 10], (d * n).sum(1), atol=1e-5)
    assert np.allclose(f[:, 11], (d * c).sum(1), atol=1e-5)
    sc = f[:, 12:60].reshape(64, 24, 2)
    assert np.allclose(sc[..., 0] ** 2 + sc[..., 1] ** 2, 1.0, atol=1e-5)
    assert np.all(f[:, 60:] == 0)
