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


def test_folded_normalisation_is_exact():
    # Training runs on standardised inputs; folding (mu, sd) into layer 1
    # must give the same network on raw inputs.
    from splishsplash_amd.learned import fold_normalisation
    rng = np.random.default_rng(11)
    th = M.init(3)
    x = rng.normal(0.0, 1.0, (64, 6)) * [9.0, 8.0, 0.1, 0.1, 4.0, 0.16] + [-10.0, -9.0, 0.05, 0.05, 1.5, 0.8]
    mu, sd = x.mean(0), x.std(0)
    y_std, _ = M.forward(th, (x - mu) / sd)
    y_raw, _ = M.forward(fold_normalisation(th, mu, sd).astype(np.float64), x)
    np.testing.assert_allclose(y_raw, y_std, rtol=2e-5, atol=2e-5)
