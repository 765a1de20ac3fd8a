"""FP64 numpy restatement of the warm-start MLP (TEST INFRASTRUCTURE).

6 -> 64 -> 64 -> 1 with SiLU, the network the device evaluates with BF16
MFMA and FP32 accumulation (splishsplash_amd/csrc/mlp.hip). Parameters are one
flat float vector in the order W1 (6x64, row-major [in][out]), b1 (64),
W2 (64x64, [in][out]), b2 (64), W3 (64), b3 (1). The loss is the mean
squared error over the batch. Pinned by central finite differences
(tests/test_oracle_mlp.py).
"""
import numpy as np

NIN, NH = 6, 64
SIZES = [("W1", (NIN, NH)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(theta):
    theta = np.asarray(theta, np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = theta[o:o + n].reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).ravel() for name, _ in SIZES])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def init(seed=0):
    rng = np.random.default_rng(seed)
    p = {"W1": rng.normal(0, np.sqrt(2.0 / NIN), (NIN, NH)), "b1": np.zeros(NH),
         "W2": rng.normal(0, np.sqrt(2.0 / NH), (NH, NH)), "b2": np.zeros(NH),
         "W3": rng.normal(0, np.sqrt(1.0 / NH), NH), "b3": np.zeros(1)}
    return pack(p)


def forward(theta, x):
    p = unpack(theta)
    x = np.asarray(x, np.float64)[:, :NIN]
    z1 = x @ p["W1"] + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = silu(z2)
    y = h2 @ p["W3"] + p["b3"][0]
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(theta, x, t):
    """Mean squared error and its gradient with respect to theta."""
    p = unpack(theta)
    y, (x, z1, h1, z2, h2) = forward(theta, x)
    n = len(y)
    r = y - np.asarray(t, np.float64)
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {"W3": h2.T @ dy, "b3": np.array([dy.sum()])}
    dz2 = np.outer(dy, p["W3"]) * dsilu(z2)
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * dsilu(z1)
    g["W1"] = x.T @ dz1
    g["b1"] = dz1.sum(0)
    return loss, pack(g)
