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
    t = np.random.default_rng(1).normal(size=(40, 6))
    p = S.init(2).astype(np.float64) + np.random.default_rng(3).normal(size=S.NPARAMS) * 0.05
    _, g = S.loss_grad(p, r, t, 100.0)
    rng = np.random.default_rng(4)
    for k in rng.choice(S.NPARAMS, 80, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (S.loss_grad(pp, r, t, 100.0)[0] - S.loss_grad(pm, r, t, 100.0)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_zero_hidden_weights_closed_form():
    # W1 = W2 = W3 = 0, zero hidden biases: y = b3 for every input,
    # dL/db3 = 2 w (b3 - mean t), and silu(0) = 0 leaves no other gradient
    p = np.zeros(S.NPARAMS)
    b3 = np.arange(6) * 0.1
    off = S.NPARAMS - 6
    p[off:] = b3
    r = sample(25, 5)
    t = np.random.default_rng(6).normal(size=(25, 6))
    y, _ = S.forward(p, r, 50.0)
    np.testing.assert_array_equal(y, np.broadcast_to(b3, y.shape))
    loss, g = S.loss_grad(p, r, t, 50.0)
    w = np.array([1, 1, 1, 2, 2, 2.0])
    np.testing.assert_allclose(g[off:], 2 * w * (b3 - t.mean(0)), rtol=1e-12)
    np.testing.assert_allclose(loss, np.mean(np.sum(w * (b3 - t) ** 2, axis=1)), rtol=1e-12)
    assert not g[:off].any()
    assert S.unpack(p)["W1"].shape == (32, 3)


def test_silu_definition():
    x = np.linspace(-30, 30, 601)
    np.testing.assert_allclose(S.silu(x), x * (1 / (1 + np.exp(-x))), rtol=1e-15)
    h = 1e-6
    np.testing.assert_allclose(S.dsilu(x), (S.silu(x + h) - S.silu(x - h)) / (2 * h), atol=1e-8)


def test_targets_and_blocks_by_hand():
    # a block with distinct off-diagonal halves: S averages them; targets
    # carry |r| = 5; blocks() rebuilds the symmetric block divided by |r|
    M = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]])
    r = np.array([[3.0, 0.0, 4.0]])
    t = S.targets(r, M.reshape(1, 9))
    np.testing.assert_allclose(t[0], 5 * np.array([1.0, 5.0, 9.0, 3.0, 7.0, 5.0]))
    B = S.blocks(r, t)[0]
    np.testing.assert_allclose(B, 0.5 * (M + M.T))
    # the loss weight makes the six-output error the block's Frobenius error
    e = np.array([[0.1, -0.2, 0.3, 0.05, -0.07, 0.02]])
    full = S.blocks(np.array([[1.0, 0, 0]]), e)[0]
    np.testing.assert_allclose(np.sum(full ** 2), np.sum(S.WEIGHT * e ** 2))
