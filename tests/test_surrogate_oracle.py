"""Pins of oracle/surrogate.py (CPU only)."""
import numpy as np

from oracle import surrogate as S


def sample(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * np.exp(rng.uniform(np.log(4), np.log(200), size=(n, 1)))


def test_gradient_matches_finite_differences():
    r = sample(40, 0)
    t = np.random.default_rng(1).normal(size=(40, 9))
    p = S.init(2).astype(np.float64) + np.random.default_rng(3).normal(size=S.NPARAMS) * 0.05
    _, g = S.loss_grad(p, r, t, 20.0)
    rng = np.random.default_rng(4)
    for k in rng.choice(S.NPARAMS, 60, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (S.loss_grad(pp, r, t, 20.0)[0] - S.loss_grad(pm, r, t, 20.0)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_zero_hidden_weights_closed_form():
    # W1 = W2 = W3 = 0: y = b3 for every input, dL/db3 = 2 mean(b3 - t) / 9
    p = np.zeros(S.NPARAMS)
    w = S.unpack(p)
    b3 = np.arange(9) * 0.1
    off = S.NPARAMS - 9
    p[off:] = b3
    r = sample(25, 5)
    t = np.random.default_rng(6).normal(size=(25, 9))
    y, _ = S.forward(p, r, 10.0)
    np.testing.assert_array_equal(y, np.broadcast_to(b3, y.shape))
    loss, g = S.loss_grad(p, r, t, 10.0)
    np.testing.assert_allclose(g[off:], 2 * (b3 - t).mean(0) / 9, rtol=1e-12)
    np.testing.assert_allclose(loss, np.mean((b3 - t) ** 2), rtol=1e-12)
    assert not g[:off].any()  # zero weights and tanh(0) = 0: no other gradient
    assert w["W1"].shape == (32, 4)


def test_features():
    r = np.array([[3.0, 0.0, 4.0], [0.0, -2.0, 0.0]])
    x = S.features(r, 5.0)
    np.testing.assert_allclose(x[0], [0.6, 0.0, 0.8, 0.0], atol=1e-15)
    np.testing.assert_allclose(x[1], [0.0, -1.0, 0.0, np.log(0.4)], atol=1e-15)
