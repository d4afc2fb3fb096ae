"""numpy float64 restatement of the far-field surrogate net, its loss and
its gradient (TEST INFRASTRUCTURE).

Only tests/ and __graft_entry__.smoke() import this module, as the checker
of the device kernels in hignn_amd/csrc/hh_surrogate.hip.  The net has no
reference counterpart (the reference evaluates every ACA entry with the full
MLP, src/FarDot.cpp:346-458 and 667-782):

    s(r) ~ |r| S(M(r)),   z = r / r_scale,
    3 -> 32 -> 32 -> 6 with SiLU hidden layers (silu x = x / (1 + e^-x)),

whose six outputs are the unique entries (xx, yy, zz, xy, yz, zx) of the
symmetrised block S(M) = M with its off-diagonal pairs averaged, times |r|.
Loss: mean over the n samples of the block's squared Frobenius error,
sum_k w_k (s_k - t_k)^2 with w = (1, 1, 1, 2, 2, 2) (each off-diagonal
entry stands for two entries of the 3x3 block).

Pinned by tests/test_surrogate_oracle.py: the gradient against central
finite differences of the loss, the closed forms of zero hidden weights
(output = b3; gradient of b3 = 2 w (b3 - mean t)), SiLU against its
definition, and the target map against a 3x3 block built by hand.
"""
import numpy as np

SHAPES = [("W1", (32, 3)), ("b1", (32,)), ("W2", (32, 32)), ("b2", (32,)), ("W3", (6, 32)),
          ("b3", (6,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 1,382
WEIGHT = np.array([1.0, 1.0, 1.0, 2.0, 2.0, 2.0])
# the six outputs as (row, col) of the 3x3 block
ENTRIES = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, off = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[off:off + k].reshape(shape)
        off += k
    return out


def init(seed):
    """Xavier-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shape[0] + shape[1]))
            parts.append(rng.uniform(-a, a, size=shape).ravel())
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def silu(x):
    return x / (1.0 + np.exp(-x))


def dsilu(x):
    s = 1.0 / (1.0 + np.exp(-x))
    return s * (1.0 + x * (1.0 - s))


def targets(r, m9):
    """|r| S(M) as the six surrogate outputs, from displacements r (n,3) and
    row-major blocks M (n,9)."""
    r = np.asarray(r, dtype=np.float64)
    M = np.asarray(m9, dtype=np.float64).reshape(-1, 3, 3)
    d = np.linalg.norm(r, axis=1)
    return np.stack([0.5 * (M[:, a, b] + M[:, b, a]) * d for a, b in ENTRIES], axis=1)


def blocks(r, s6):
    """Symmetric 3x3 blocks s / |r| (n,3,3) from surrogate outputs."""
    r = np.asarray(r, dtype=np.float64)
    d = np.linalg.norm(r, axis=1)
    out = np.zeros((r.shape[0], 3, 3))
    for k, (a, b) in enumerate(ENTRIES):
        out[:, a, b] = out[:, b, a] = s6[:, k] / d
    return out


def forward(p, r, r_scale):
    w = unpack(p)
    z = np.asarray(r, dtype=np.float64) / r_scale
    x1 = z @ w["W1"].T + w["b1"]
    a1 = silu(x1)
    x2 = a1 @ w["W2"].T + w["b2"]
    a2 = silu(x2)
    y = a2 @ w["W3"].T + w["b3"]
    return y, (z, x1, a1, x2, a2)


def loss_grad(p, r, t, r_scale):
    """(loss, gradient as a flat vector in the parameter order)."""
    w = unpack(p)
    y, (z, x1, a1, x2, a2) = forward(p, r, r_scale)
    n = y.shape[0]
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(WEIGHT * e * e) / n)
    dy = 2.0 * WEIGHT * e / n
    g = {"W3": dy.T @ a2, "b3": dy.sum(0)}
    d2 = (dy @ w["W3"]) * dsilu(x2)
    g["W2"], g["b2"] = d2.T @ a1, d2.sum(0)
    d1 = (d2 @ w["W2"]) * dsilu(x1)
    g["W1"], g["b1"] = d1.T @ z, d1.sum(0)
    return loss, np.concatenate([g[name].ravel() for name, _ in SHAPES])
