"""The far-field surrogate's own GPU kernels against oracle/surrogate.py, and
what the trained network gains over the full MLP (GPU)."""
import numpy as np
import pytest

from hignn_amd import abi, weights
from oracle import surrogate as S

pytestmark = pytest.mark.gpu

R0 = 20.0


def far_pairs(n, seed, lo=4.0, hi=200.0):
    """Displacements with |r| log-uniform in [lo, hi] and uniform directions
    (the far-field range of the blob workloads)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * np.exp(rng.uniform(np.log(lo), np.log(hi), size=(n, 1)))).astype(np.float32)


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.device("cuda:0"))


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_gradient_and_forward_match_oracle():
    r = far_pairs(3000, 0)
    t = np.random.default_rng(1).normal(size=(3000, 9)).astype(np.float32)
    p = (S.init(2) + np.random.default_rng(3).normal(size=S.NPARAMS) * 0.05).astype(np.float32)
    s = abi.Surrogate(R0)
    assert not s.params().any()  # a new surrogate starts from zero parameters
    s.set_params(p)
    np.testing.assert_array_equal(s.params(), p)
    rd, td = dev(r), dev(t)
    loss, g = s.gradient(rd.data_ptr(), td.data_ptr(), r.shape[0])
    loss64, g64 = S.loss_grad(p, r, t, R0)
    print(f"loss {loss:.6e} vs {loss64:.6e}; gradient rel L2 {rel(g, g64):.2e}")
    assert abs(loss - loss64) <= 1e-5 * loss64
    assert rel(g, g64) <= 1e-5
    # zero steps: the loss at the current parameters, parameters unchanged
    # (the loss is an atomic sum over workgroups: equal up to summation order)
    assert abs(s.train(rd.data_ptr(), td.data_ptr(), r.shape[0], 0, 1e-3) - loss) <= 1e-6 * loss
    np.testing.assert_array_equal(s.params(), p)
    out = dev(np.zeros((3000, 9)))
    s.eval(rd.data_ptr(), r.shape[0], out.data_ptr())
    y64, _ = S.forward(p, r, R0)
    assert rel(out.cpu().numpy(), y64) <= 1e-6


def test_training_and_gain_over_full_mlp():
    """Train on pairs labelled by the full MLP (|r| M(r)) and measure, on
    held-out pairs, what the surrogate gives: its error against the full MLP,
    and its cost per entry against the full MLP's."""
    import torch
    w = weights.xavier(0)
    m = abi.Model(far_pairs(64, 9), 16)  # any model: eval_pair only needs the weights
    m.load_weights(w)
    r_tr, r_te = far_pairs(1 << 16, 4), far_pairs(1 << 14, 5)
    lab = lambda r: (m.eval_pair(r, "fp32") * np.linalg.norm(r, axis=1, keepdims=True)).astype(np.float32)
    t_tr, t_te = lab(r_tr), lab(r_te)
    s = abi.Surrogate(R0)
    s.set_params(S.init(7))
    rd, td = dev(r_tr), dev(t_tr)
    loss0, _ = s.gradient(rd.data_ptr(), td.data_ptr(), r_tr.shape[0])
    loss = s.train(rd.data_ptr(), td.data_ptr(), r_tr.shape[0], 3000, 3e-3)
    out = dev(np.zeros((r_te.shape[0], 9)))
    s.eval(dev(r_te).data_ptr(), r_te.shape[0], out.data_ptr())
    err = rel(out.cpu().numpy(), t_te)
    # cost per entry: FLOP count and measured device time for 2^22 entries
    big = dev(far_pairs(1 << 22, 6))
    yb = dev(np.zeros((1 << 22, 9)))
    s.eval(big.data_ptr(), 1 << 22, yb.data_ptr())  # warm-up
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    s.eval(big.data_ptr(), 1 << 22, yb.data_ptr())
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1)
    flop_s = 2 * (4 * 32 + 32 * 32 + 32 * 9)
    print(f"surrogate: loss {loss0:.3e} -> {loss:.3e}; held-out error vs full MLP {err:.3e}; "
          f"{flop_s} FLOP per entry against 141056 ({141056 / flop_s:.0f}x fewer); "
          f"{(1 << 22) / ms / 1e6:.2f} G entries/s")
    assert loss < 0.1 * loss0  # the kernels train the net
    assert err < 0.5  # it learns the entries' bulk
    assert 141056 / flop_s > 40
