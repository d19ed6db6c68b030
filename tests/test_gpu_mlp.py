"""GPU tests of the warm-start MLP kernels.

* loss and every parameter gradient of the HIP forward/backward pass agree
  with the CPU oracle (oracle/mlp.py, BF16 rounding at the same points;
  tolerance 2e-2 relative Frobenius per layer block for the FP32 accumulation
  order and the BF16 flips it causes);
* predictions agree with the oracle;
* training on a smooth synthetic target lowers the loss by 10x;
* on a small frequency sweep the trained network's guesses lower the mean
  GMRES iteration count on held-out wavenumbers, with every solve converged
  on the recomputed FP64 residual (the learned-warm-start gain).
"""

import numpy as np
import pytest

import sonium_amd as sa
from oracle import mlp as omlp
from sonium_amd import warmstart as ws
from sonium_amd.mesh import icosphere

from helpers import rel_fro

pytestmark = pytest.mark.gpu


def _raw(rng, m):
    c = rng.standard_normal((m, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    k = rng.uniform(1, 12, size=(m, 1))
    return np.concatenate([c, c, d, k], axis=1).astype(np.float32)


def _blocks(width):
    off = 0
    for rows, cols in omlp.layer_shapes(width):
        yield off, rows * cols
        off += rows * cols
        yield off, rows
        off += rows


def test_gradients_match_oracle(ctx):
    net = ws.WarmStartMLP(ctx, width=512, seed=11)
    rng = np.random.default_rng(1)
    raw = _raw(rng, 512)
    t = rng.standard_normal((512, 2)).astype(np.float32)
    loss, g = net.loss_and_gradients(raw, t)
    p = net.params().astype(np.float64)
    loss_ref, g_ref, _ = omlp.loss_and_gradients(p, 512, raw, t, bf16_storage=True)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
    for off, cnt in _blocks(512):
        a, b = g[off: off + cnt], g_ref[off: off + cnt]
        assert rel_fro(a, b) < 2e-2, (off, cnt, rel_fro(a, b))


def test_predictions_match_oracle(ctx):
    net = ws.WarmStartMLP(ctx, width=512, seed=12)
    rng = np.random.default_rng(2)
    raw = _raw(rng, 1000)
    y = net.predict(raw)
    y_ref = omlp.predict(net.params().astype(np.float64), 512, raw)
    assert rel_fro(y, y_ref) < 1e-2


@pytest.mark.parametrize("width", [64, 512, 1024, 2048])
def test_fused_inference_equals_layered_forward(ctx, width):
    """The fused inference kernel (all layers per 32-sample workgroup,
    activations in LDS) gives the layer-by-layer forward pass of training bit
    for bit, on ragged sample counts and after training steps; widths above
    1024 take the layered path."""
    net = ws.WarmStartMLP(ctx, width=width, seed=21)
    rng = np.random.default_rng(4)
    for m in (1, 257, 3000):
        raw = _raw(rng, m)
        assert np.array_equal(net.predict(raw), net.predict_layered(raw))
    net.set_dataset(_raw(rng, 4096), rng.standard_normal((4096, 2)).astype(np.float32))
    net.train(5, batch=1024)
    raw = _raw(rng, 777)
    assert np.array_equal(net.predict(raw), net.predict_layered(raw))
    net.close()


def test_training_lowers_loss(ctx):
    net = ws.WarmStartMLP(ctx, width=512, seed=13)
    rng = np.random.default_rng(3)
    raw = _raw(rng, 1 << 15)
    dc = (raw[:, 0:3] * raw[:, 6:9]).sum(1)
    k = raw[:, 9]
    t = np.stack([np.cos(0.3 * k * dc), np.sin(0.2 * k) * dc], axis=1).astype(np.float32)
    net.set_dataset(raw, t)
    losses = net.train(600, batch=4096, lr_max=2e-3, lr_min=1e-4, seed=3)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.1 * losses[:5].mean(), (losses[:5], losses[-20:])


def test_learned_warm_start_lowers_iterations(ctx):
    m = icosphere(1.0, 3)
    tm = sa.TbemMesh(ctx, m)
    c, n, _ = tm.geometry()
    wsp = sa.TbemAssemblyWorkspace(ctx, tm)
    d = np.array([0.0, 0.0, 1.0])
    ks = np.linspace(1.0, 4.0, 49)
    cfg = sa.GmresConfig(max_iterations=10, restart=100, tolerance=1e-8)
    sols, inputs, targets = {}, [], []
    for i, k in enumerate(ks):
        ph = sa.PhysicsParams.from_wavenumber(k)
        beta = ph.burton_miller_beta()
        system = wsp.assemble(ph, beta)
        b = sa.IncidentField([d]).compute_rhs_with_beta(tm, ph, beta)
        op = sa.DenseMatrixOperator(system)
        sol = sa.gmres(op, b, cfg)
        assert sol.converged
        sols[i] = (op, b, sol)
        if i % 2 == 0:
            inputs.append(ws.panel_inputs(c, n, k, d))
            targets.append(ws.density_to_target(sol.x, c, k, d))
    net = ws.WarmStartMLP(ctx, width=512, seed=5)
    net.set_dataset(np.concatenate(inputs), np.concatenate(targets))
    losses = net.train(800, batch=8192, lr_max=2e-3, lr_min=5e-5, seed=5)
    zero, learned = [], []
    for i in range(1, len(ks), 2):
        op, b, sol = sols[i]
        x0 = net.guess(c, n, ks[i], d)
        s2, used = ws.solve_with_guess(op, b, x0, cfg)
        assert s2.converged and s2.residual < 1e-8
        assert np.linalg.norm(b - op.apply(s2.x)) / np.linalg.norm(b) < 1e-8
        zero.append(sol.iterations)
        learned.append(s2.iterations)
    print("final loss", losses[-1], "zero-guess iterations", np.mean(zero), "learned", np.mean(learned))
    assert np.mean(learned) < np.mean(zero)
