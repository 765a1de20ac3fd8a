"""BF16 MFMA kernels of the warm-start network against the FP64 numpy
network. Inputs, weights and activations are rounded to BF16 on the device
(FP32 accumulation), so agreement is normwise at the BF16 level."""
import numpy as np
import pytest

from oracle import mlp_oracle as M
from splishsplash_amd import capi

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("n", [1, 31, 32, 33, 1000, 100003])
def test_forward_matches_oracle(n):
    rng = np.random.default_rng(n)
    theta = M.init(3).astype(np.float32)
    x = rng.normal(size=(n, 6)).astype(np.float32)
    y = capi.mlp_forward(theta, x)
    # oracle with the BF16 rounding of inputs and weights (activations stay FP64)
    th = theta.astype(np.float64).copy()
    p = M.unpack(th)
    for k in ("W1", "W2", "W3"):
        p[k][...] = bf16(p[k])
    yo, _ = M.forward(M.pack(p), bf16(x))
    assert rel(y, yo) < 2e-2
    assert np.corrcoef(y, yo)[0, 1] > 0.999 if n > 2 else True


def test_gradient_matches_oracle():
    rng = np.random.default_rng(7)
    theta = M.init(4).astype(np.float32)
    x = rng.normal(size=(5000, 6)).astype(np.float32)
    t = rng.normal(size=5000).astype(np.float32)
    loss, g = capi.mlp_loss_grad(theta, x, t)
    lo, go = M.loss_and_grad(theta.astype(np.float64), x.astype(np.float64), t)
    assert loss == pytest.approx(lo, rel=3e-2)
    for name, o, sz in [("W1", 0, 384), ("b1", 384, 64), ("W2", 448, 4096), ("b2", 4544, 64), ("W3", 4608, 64),
                        ("b3", 4672, 1)]:
        e = rel(g[o:o + sz], go[o:o + sz])
        assert e < 5e-2, (name, e)
    cos = g @ go / (np.linalg.norm(g) * np.linalg.norm(go))
    assert cos > 0.999


def test_training_fits_a_smooth_target():
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, size=(20000, 6)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] + 0.5 * x[:, 3] ** 2).astype(np.float32)
    theta = M.init(5).astype(np.float32)
    l0 = float(np.mean((M.forward(theta, x)[0] - t) ** 2))
    th, l1 = capi.mlp_train(theta, x, t, steps=1500, batch=2048, lr=3e-3, seed=1)
    assert l1 < 0.05 * l0, (l0, l1)
    yo, _ = M.forward(th.astype(np.float64), x)
    assert float(np.mean((yo - t) ** 2)) < 0.1 * l0
