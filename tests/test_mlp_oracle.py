"""Pins of the CPU restatement of the learned gain network (oracle/mlp_oracle.py):
bf16 rounding against hand-computed values, the float64 backward pass
against central finite differences, Adam's first step against its closed
form, and the group features against their definition."""
import numpy as np

from oracle import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, np.pi, -2.5e-3, 65504.0, 2 ** -130],
                 np.float32)
    # 1+2^-8 is halfway between 1 and 1+2^-7: ties to even -> 1
    want = [1.0, 1.0, 1.0 + 2 ** -7, 3.140625, -0.0025024414, 65536.0, 2 ** -130]
    got = M.bf16_round(x)
    np.testing.assert_allclose(got[:5], want[:5], rtol=0, atol=1e-9)
    assert got[5] == 65536.0
    assert got[6] == np.float32(2 ** -130)   # subnormal with 7 mantissa bits is exact


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = M.init_params(5).astype(np.float64)
    x = rng.normal(size=(9, M.N_IN))
    t = rng.uniform(size=(9, M.N_OUT))
    _, g = M.backward(theta, x, t, rounding=False)
    eps = 1e-6
    for i in rng.choice(M.N_PARAMS, 40, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fd = (M.loss(M.forward(tp, x, False)[0], t) - M.loss(M.forward(tm, x, False)[0], t)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-4, abs(fd)) + 1e-12, (i, fd, g[i])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(0)
    theta = rng.normal(size=50)
    g = rng.normal(size=50)
    lr = 1e-3
    t1, m, v = M.adam(theta, np.zeros(50), np.zeros(50), g, 1, lr)
    # bias-corrected moments of the first step are g and g^2
    np.testing.assert_allclose(t1, theta - lr * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_features_definition():
    rng = np.random.default_rng(1)
    sigma, N = 20.0, 32
    z = rng.normal(0, sigma, (64, N))
    x = M.features(z, N, sigma)
    assert x.shape == (N, 66)
    np.testing.assert_allclose(x[:, :64], z.T / sigma)
    assert np.all(x[:, 64] == 0.0)                      # log(20/20)
    e = np.array([(z[:, n] ** 2).sum() for n in range(N)]) / (64 * sigma ** 2)
    np.testing.assert_allclose(x[:, 65], np.log1p(e))
    assert 0.6 < e.mean() < 1.4                           # ~1 for pure noise
