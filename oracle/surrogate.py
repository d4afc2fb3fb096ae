"""numpy float64 restatement of the far-field surrogate net, its loss and
its gradient (TEST INFRASTRUCTURE).

Only tests/ import this module, as the checker of the device kernels in
hignn_amd/csrc/hh_surrogate.hip.  The net has no reference counterpart (the
reference evaluates every ACA entry with the full MLP, src/FarDot.cpp:330-470):
s(r) ~ |r| M(r), inputs (r/|r|, ln(|r|/r0)), 4 -> 32 -> 32 -> 9 with tanh
hidden layers; loss = mean squared error over the 9 outputs of n samples.

Pinned by tests/test_surrogate_oracle.py: the gradient against central
finite differences of the loss, the closed forms of zero hidden weights
(output = b3, gradient of b3 = 2 mean error / 9) and the feature map.
"""
import numpy as np

SHAPES = [("W1", (32, 4)), ("b1", (32,)), ("W2", (32, 32)), ("b2", (32,)), ("W3", (9, 32)),
          ("b3", (9,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, off = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[off:off + k].reshape(shape)
        off += k
    return out


def init(seed):
    """Xavier-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-a, a, size=shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def features(r, r0):
    r = np.asarray(r, dtype=np.float64)
    d = np.linalg.norm(r, axis=1, keepdims=True)
    return np.concatenate([r / d, np.log(d / r0)], axis=1)


def forward(p, r, r0):
    w = unpack(p)
    x = features(r, r0)
    a1 = np.tanh(x @ w["W1"].T + w["b1"])
    a2 = np.tanh(a1 @ w["W2"].T + w["b2"])
    y = a2 @ w["W3"].T + w["b3"]
    return y, (x, a1, a2)


def loss_grad(p, r, t, r0):
    """(loss, gradient as a flat vector in the parameter order)."""
    w = unpack(p)
    y, (x, a1, a2) = forward(p, r, r0)
    n = y.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(e * e) / (9 * n))
    dy = 2.0 * e / (9 * n)
    g = {"W3": dy.T @ a2, "b3": dy.sum(0)}
    d2 = (dy @ w["W3"]) * (1 - a2 * a2)
    g["W2"], g["b2"] = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ w["W2"]) * (1 - a1 * a1)
    g["W1"], g["b1"] = d1.T @ x, d1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])
