"""Oracle kernels against the reference's compiled SPHKernels.cpp (golden
probe) and the closed forms of Tests/Kernel/KernelTests.cpp:17-47."""
import numpy as np
import pytest

from oracle import oracle


@pytest.mark.parametrize("ki", [0, 1, 2])
def test_table_matches_reference_build(golden, ki):
    g = golden["kernels"][ki]
    T, G, s = oracle.kernel_table(g["h"])
    idx = np.asarray(g["index"], dtype=int)
    np.testing.assert_array_equal(T[idx], np.asarray(g["T"]))
    np.testing.assert_array_equal(G[idx], np.asarray(g["G"]))
    assert s["W_zero"] == g["W_zero"]
    assert s["invStep"] == g["invStep"]
    assert abs(T.sum() - g["sumT"]) <= 1e-13 * abs(g["sumT"])
    assert abs(G.sum() - g["sumG"]) <= 1e-13 * abs(g["sumG"])


@pytest.mark.parametrize("ki", [0, 1, 2])
def test_kernel_values_match_reference_build(golden, ki):
    g = golden["kernels"][ki]
    r = np.asarray(g["r"]).reshape(-1, 3)
    w, gw = oracle.kernel_eval(g["h"], r, precomputed=True)
    np.testing.assert_allclose(w, g["W"], rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(gw.ravel(), g["gradW"], rtol=1e-13, atol=1e-12 * np.abs(g["gradW"]).max())
    wc, gc = oracle.kernel_eval(g["h"], r, precomputed=False)
    np.testing.assert_allclose(wc, g["W_cubic"], rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(gc.ravel(), g["gradW_cubic"], rtol=1e-12,
                               atol=1e-12 * np.abs(g["gradW_cubic"]).max())


@pytest.mark.parametrize("precomputed", [True, False])
def test_kernel_normalised_positive(precomputed):
    # KernelTests.cpp:17-47: 50^3 grid over [-h, h]^3, h = 0.1, eps 1e-5 (FP64)
    h = 0.1
    n = 50
    step = 2 * h / (n - 1)
    a = -h + step * np.arange(n)
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    r = -np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    w, g = oracle.kernel_eval(h, r, precomputed=precomputed)
    V = step ** 3
    assert abs(w.sum() * V - 1.0) < 1e-5
    assert np.linalg.norm(g.sum(axis=0) * V) < 1e-5
    assert (w >= -1e-5).all()


def test_table_is_piecewise_constant_midpoint():
    # W(r) = 0.5 (T[p] + T[p+1]) with p = floor(r / step): constant inside a
    # step, and the table samples the analytic cubic spline at p * step.
    h = 0.1
    T, G, s = oracle.kernel_table(h)
    step = h / 9999
    p = np.array([0, 1, 17, 4999, 5000, 9000, 9997])
    r = np.zeros((len(p) * 2, 3))
    r[0::2, 0] = (p + 0.25) * step
    r[1::2, 0] = (p + 0.75) * step
    w, _ = oracle.kernel_eval(h, r)
    np.testing.assert_array_equal(w[0::2], w[1::2])
    np.testing.assert_allclose(w[0::2], 0.5 * (T[p] + T[p + 1]), rtol=0, atol=0)
    wc, _ = oracle.kernel_eval(h, np.stack([p * step, 0 * p, 0 * p], axis=1), precomputed=False)
    np.testing.assert_allclose(T[p], wc, rtol=1e-14)
    k = 8 / (np.pi * h ** 3)
    assert s["W_zero"] == pytest.approx(k, rel=1e-15)
