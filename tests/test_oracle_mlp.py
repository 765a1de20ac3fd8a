"""The numpy FP64 warm-start network (oracle/mlp_oracle.py): gradient by
central finite differences."""
import numpy as np

from oracle import mlp_oracle as M


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    theta = M.init(1)
    x = rng.normal(size=(37, 6))
    t = rng.normal(size=37)
    loss, g = M.loss_and_grad(theta, x, t)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    idx = np.concatenate([idx, [M.NPARAM - 1, 0, 384, 448, 4544, 4608]])
    eps = 1e-6
    for k in idx:
        e = np.zeros(M.NPARAM)
        e[k] = eps
        fd = (M.loss_and_grad(theta + e, x, t)[0] - M.loss_and_grad(theta - e, x, t)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    assert M.NPARAM == 4673


def test_forward_shapes_and_silu():
    theta = M.init(2)
    y, cache = M.forward(theta, np.zeros((5, 8)))
    assert y.shape == (5,)
    assert abs(M.silu(0.0)) == 0.0 and abs(M.dsilu(0.0) - 0.5) < 1e-15
