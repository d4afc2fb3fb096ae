"""The far-field surrogate on the GPU: its kernels against oracle/surrogate.py,
training on samples labelled by the full MLP, and the judged surrogate ACA
matvec against the FP64 evaluation."""
import numpy as np
import pytest

from hignn_amd import abi, weights, workloads
from oracle import oracle as O
from oracle import surrogate as S

pytestmark = pytest.mark.gpu

R_LO, R_HI = 4.0, 200.0


def far_pairs(n, seed, lo=R_LO, hi=R_HI):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * np.exp(rng.uniform(np.log(lo), np.log(hi), size=(n, 1)))).astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def test_gradient_and_forward_match_oracle():
    r = far_pairs(3000, 0)
    t = np.random.default_rng(1).normal(size=(3000, 6)).astype(np.float32)
    p = (S.init(2) + np.random.default_rng(3).normal(size=S.NPARAMS) * 0.05).astype(np.float32)
    s = abi.Surrogate(R_LO, R_HI)
    assert not s.params().any()  # a new surrogate starts from zero parameters
    s.set_params(p)
    np.testing.assert_array_equal(s.params(), p)
    rd, td = abi.DeviceArray.of(r), abi.DeviceArray.of(t)
    loss, g = s.gradient(rd.ptr, td.ptr, r.shape[0])
    loss64, g64 = S.loss_grad(p, r, t, R_HI)
    print(f"loss {loss:.6e} vs {loss64:.6e}; gradient rel L2 {rel(g, g64):.2e}")
    assert abs(loss - loss64) <= 1e-5 * loss64
    assert rel(g, g64) <= 1e-5
    # the gradient adds block partials in a fixed order: bit-identical repeats
    loss2, g2 = s.gradient(rd.ptr, td.ptr, r.shape[0])
    assert loss2 == loss and np.array_equal(g2, g)
    # zero steps: the loss at the current parameters, parameters unchanged
    assert s.train(rd.ptr, td.ptr, r.shape[0], 0, 1e-3) == loss
    np.testing.assert_array_equal(s.params(), p)
    out = abi.DeviceArray((3000, 6))
    s.eval(rd.ptr, r.shape[0], out.ptr)
    y64, _ = S.forward(p, r, R_HI)
    assert rel(out.numpy(), y64) <= 2e-6


def test_samples_are_labelled_by_the_full_mlp():
    w = weights.smooth(0)
    m = abi.Model(workloads.uniform_cube(200, 0.05, 0), 64)
    m.load_weights(w)
    n = 20000
    rd, td = m.surrogate_samples(n, R_LO, R_HI, seed=7)
    r, t = rd.numpy(), td.numpy()
    d = np.linalg.norm(r.astype(np.float64), axis=1)
    assert d.min() >= R_LO * (1 - 1e-6) and d.max() <= R_HI * (1 + 1e-6)
    # log-uniform radii: quartiles of ln(d) at a quarter of the range each
    q = np.quantile(np.log(d / R_LO) / np.log(R_HI / R_LO), [0.25, 0.5, 0.75])
    np.testing.assert_allclose(q, [0.25, 0.5, 0.75], atol=0.02)
    # uniform directions: mean direction near zero, second moment I/3
    u = r / d[:, None]
    assert np.abs(u.mean(0)).max() < 0.03
    np.testing.assert_allclose(u.T @ u / n, np.eye(3) / 3, atol=0.02)
    # targets |r| S(M) from the FP32 pair function (oracle's float MLP)
    want = S.targets(r, O.mlp_f32(w, r))
    assert rel(t, want) <= 1e-5
    # a different seed gives different samples, the same seed the same ones
    r2 = m.surrogate_samples(1000, R_LO, R_HI, seed=7)[0].numpy()
    np.testing.assert_array_equal(r2, r[:1000])
    assert not np.array_equal(m.surrogate_samples(1000, R_LO, R_HI, seed=8)[0].numpy(), r2)


@pytest.fixture(scope="module")
def trained():
    """A surrogate of the smooth weight set, trained on the GPU on 2^21
    samples labelled by the full MLP (|r| in [4, 200])."""
    w = weights.smooth(0)
    m = abi.Model(workloads.uniform_cube(200, 0.05, 0), 64)
    m.load_weights(w)
    n = 1 << 21
    rd, td = m.surrogate_samples(n, R_LO, R_HI, seed=1)
    s = abi.Surrogate(R_LO, R_HI)
    s.set_params(S.init(0))
    first = s.train(rd.ptr, td.ptr, n, 0, 1e-3)
    last = s.train(rd.ptr, td.ptr, n, 20000, 3e-3, batch=16384, lr_final=3e-5)
    return w, s, first, last


def test_training_and_gain_over_full_mlp(trained):
    """Held-out error of the trained surrogate against the full MLP, and its
    cost per entry against the full MLP's."""
    w, s, first, last = trained
    r = far_pairs(1 << 16, 123)
    out = abi.DeviceArray((r.shape[0], 6))
    s.eval(abi.DeviceArray.of(r).ptr, r.shape[0], out.ptr)
    y = out.numpy().astype(np.float64)
    B = S.blocks(r, y)
    M = O.mlp_f32(w, r).astype(np.float64).reshape(-1, 3, 3)
    full = 0.5 * (M + M.transpose(0, 2, 1))
    e = np.linalg.norm(B - full, axis=(1, 2)) / np.linalg.norm(full, axis=(1, 2))
    med, p90 = np.median(e), np.quantile(e, 0.9)
    print(f"loss {first:.3e} -> {last:.3e}; held-out block error median {med:.2e} p90 {p90:.2e}")
    assert last < 1e-3 * first
    # the judge's default bar is 2e-3 per block: most entries are well inside
    assert med < 1e-3 and p90 < 4e-3
    # FLOP per entry: 2 (3*32 + 32*32 + 32*6) against 2 (3*128 + 128*512 + 512*9)
    assert 2 * (3 * 32 + 32 * 32 + 32 * 6) * 50 < 141056


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_judged_surrogate_matvec_vs_fp64(trained, prec):
    """The ACA on surrogate entries for the blocks that pass the judge and on
    full-MLP entries for the rest: velocities against the FP64 evaluation
    of the same blocks, with counts of judged and fallen-back blocks."""
    w, s, _, _ = trained
    x = workloads.uniform_cube(12000, 0.05, 3)
    f = workloads.gravity(x.shape[0])
    m = abi.Model(x, 64)
    m.load_weights(w)
    m.set(epsilon=1e-4, max_iter=40, precision=prec)
    m.update_coord()
    tg = np.arange(0, x.shape[0], 11)
    ref = m.check_fp64(f, tg)
    u_full = m.dot(f)
    t_full = m.timings()[1]
    m.set_surrogate(s, tol=2e-3)
    u_sur = m.dot(f)
    t_sur = m.timings()[1]
    judged, on_sur, back, n_sur, n_full = m.surrogate_stats()
    e_full, e_sur = rel(u_full[tg], ref), rel(u_sur[tg], ref)
    print(f"{prec}: far blocks judged {judged}, on surrogate {on_sur}, fell back {back}; "
          f"entries surrogate {n_sur} full {n_full}; far ms {t_full:.2f} -> {t_sur:.2f}; "
          f"error vs FP64 full {e_full:.2e} surrogate {e_sur:.2e}")
    assert judged == len(m.far_ranks()) and on_sur + back == judged
    assert on_sur > 0.5 * judged
    # the blocks on surrogate entries are exactly those judged within the bar
    errs = m.surrogate_errors()
    assert int(np.sum(errs <= 2e-3)) == on_sur
    assert e_sur <= 1e-3
    # on one near workgroup the 16-bit path runs the rejected blocks' loop
    # first, beside the near kernel: the same ranks, judge split and
    # velocities up to the order of the FP64 velocity sums (a bar at the
    # median judged error sends half the blocks to the full MLP)
    m.set_surrogate(s, tol=float(np.median(errs)))
    u_a = m.dot(f)
    ranks_a, stats_a = m.far_ranks(), tuple(m.surrogate_stats()[:3])
    m.set_near_workgroups(1)
    u_cap = m.dot(f)
    m.set_near_workgroups(0)
    assert stats_a[2] > 0.3 * judged
    np.testing.assert_array_equal(m.far_ranks(), ranks_a)
    assert tuple(m.surrogate_stats()[:3]) == stats_a
    assert rel(u_cap, u_a) <= 1e-6
    # the surrogate is detached: the next matvec is the full-MLP one again
    m.set_surrogate(None)
    np.testing.assert_allclose(m.dot(f), u_full, rtol=0, atol=1e-6 * np.abs(u_full).max())
