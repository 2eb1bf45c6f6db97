"""CPU float64 restatement of the learned gain network (test infrastructure only).

The network has no counterpart in the reference; it is the learned single
pass that replaces the second step: per group row, 66 features -> 128 -> 128 -> 64 gains,
ReLU hidden layers, sigmoid output, weights stored as bfloat16 with float32
master copies, trained with Adam on the mean squared error against the
step-2 Wiener gains y^2 / (y^2 + (beta sigma)^2) (src/core/nldct.cpp:2131)
that the plain two-step method produces.

Forward semantics shared with the HIP kernels (bm4d-of_amd/csrc/mlp.hip):
  * the weights used by the forward pass are the master weights rounded to
    bfloat16 (round to nearest even); biases stay float32;
  * the layer inputs (features and both hidden activations) are rounded to
    bfloat16 before each product, as the bf16 matrix cores consume them;
  * backward is straight-through: the rounding has derivative 1.
This module computes all of that in float64 from the rounded values, so the
GPU's float32-accumulated results must agree to float32 accumulation error.

Only tests/ load this module.
"""
import numpy as np

N_IN, N_H, N_OUT = 66, 128, 64
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_OUT)), ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_round(x):
    """float -> nearest bfloat16 (ties to even), returned as float64.  NaN
    and infinities are not expected here."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(theta):
    """Flat float32 parameter vector (layout of the HIP master copy) ->
    dict of arrays.  W[k][j] is stored k-major (input index outermost)."""
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shp)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[n], np.float64).ravel() for n, _ in SHAPES])


def init_params(seed=0):
    """He-style initialisation used by the training driver."""
    rng = np.random.default_rng(seed)
    d = {}
    d["W1"] = rng.normal(0, np.sqrt(2.0 / N_IN), (N_IN, N_H))
    d["b1"] = np.zeros(N_H)
    d["W2"] = rng.normal(0, np.sqrt(2.0 / N_H), (N_H, N_H))
    d["b2"] = np.zeros(N_H)
    d["W3"] = rng.normal(0, np.sqrt(1.0 / N_H), (N_H, N_OUT))
    d["b3"] = np.zeros(N_OUT)
    return pack(d).astype(np.float32)


def forward(theta, x, rounding=True):
    """x: (B, 66).  Returns (gains (B, 64), cache for backward)."""
    p = unpack(theta)
    rd = bf16_round if rounding else (lambda a: np.asarray(a, np.float64))
    W1, W2, W3 = rd(p["W1"]), rd(p["W2"]), rd(p["W3"])
    x0 = rd(x)
    z1 = x0 @ W1 + p["b1"]
    h1 = np.maximum(z1, 0.0)
    h1r = rd(h1)
    z2 = h1r @ W2 + p["b2"]
    h2 = np.maximum(z2, 0.0)
    h2r = rd(h2)
    z3 = h2r @ W3 + p["b3"]
    g = 1.0 / (1.0 + np.exp(-z3))
    return g, (x0, z1, h1r, z2, h2r, g, W1, W2, W3)


def loss(g, t):
    """Mean squared error over all B*64 gains."""
    return float(np.mean((g - t) ** 2))


def backward(theta, x, t, rounding=True):
    """Gradient of loss(forward(theta, x), t) w.r.t. the flat master
    parameters (straight-through rounding).  Returns (loss, grad)."""
    g, (x0, z1, h1r, z2, h2r, _, W1, W2, W3) = forward(theta, x, rounding)
    B = x.shape[0]
    dg = 2.0 * (g - t) / (B * N_OUT)
    dz3 = dg * g * (1.0 - g)
    dW3 = h2r.T @ dz3
    db3 = dz3.sum(0)
    dh2 = dz3 @ W3.T
    dz2 = dh2 * (z2 > 0)
    dW2 = h1r.T @ dz2
    db2 = dz2.sum(0)
    dh1 = dz2 @ W2.T
    dz1 = dh1 * (z1 > 0)
    dW1 = x0.T @ dz1
    db1 = dz1.sum(0)
    grad = pack({"W1": dW1, "b1": db1, "W2": dW2, "b2": db2, "W3": dW3, "b3": db3})
    return loss(g, t), grad


def adam(theta, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba, Algorithm 1) with bias correction;
    `step` counts from 1.  Returns new (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def features(z, N, sigma):
    """Features of the N rows of one group from its Haar-domain step-1
    coefficients z[k][n] (64 x N): z/sigma, log(sigma/20) and the row
    statistic log1p(sum_k z[k][n]^2 / (64 sigma^2)) (for a detail row the
    squared differences of the patch sub-means it compares: a distance)."""
    z = np.asarray(z, np.float64)[:, :N]
    x = np.zeros((N, N_IN))
    x[:, :64] = (z / sigma).T
    x[:, 64] = np.log(sigma / 20.0)
    x[:, 65] = np.log1p((z ** 2).sum(0) / (64.0 * sigma * sigma))
    return x
