"""GPU parity of libhh against the CPU oracle, called through the C ABI.

Bars: FP32 path bit-exact lists and <= 1e-5 relative L2 velocity error
against the oracle; FP64 check path <= 1e-10 against the oracle's FP64
evaluation; 16-bit forms <= 1e-3 against the FP64 block-convention matvec.
"""
import numpy as np
import pytest

from hignn_amd import abi, weights, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _disp(n, seed):
    g = np.random.default_rng(seed)
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.exp(g.uniform(np.log(1e-6), np.log(1e3), size=n))
    out = (d * r[:, None]).astype(np.float32)
    out[0] = 0.0
    out[1] = [3e-6, 0, 0]  # below the 1e-5 identity cut-off
    return out


@pytest.fixture(scope="module")
def model_small(xavier_w):
    x = workloads.uniform_cube(1200, 0.05, 0)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50)
    m.update_coord()
    om = O.OracleModel(x, 64, epsilon=1e-5, max_iter=50)
    om.update_coord()
    f = workloads.gravity(x.shape[0])
    return dict(x=x, m=m, om=om, f=f)


def test_pair_fp32_matches_oracle(xavier_w):
    x = workloads.uniform_cube(100, 0.05, 0)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    r = _disp(4099, 1)
    got = m.eval_pair(r, "fp32").astype(np.float64)
    ref = O.mlp_f64(xavier_w, r.astype(np.float64))
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(got - ref) / scale
    print("fp32 pair max rel err", err.max())
    assert err.max() < 2e-5
    np.testing.assert_array_equal(got[0], np.eye(3).reshape(-1))
    np.testing.assert_array_equal(got[1], np.eye(3).reshape(-1))


def test_pair_fp64_matches_oracle(xavier_w):
    x = workloads.uniform_cube(100, 0.05, 0)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    r = _disp(1000, 2)
    got = m.eval_pair(r, "fp64").astype(np.float64)
    ref = O.mlp_f64(xavier_w, r.astype(np.float64))
    # output is rounded to float32
    err = np.abs(got - ref) / np.abs(ref).max(axis=1, keepdims=True)
    print("fp64 pair max rel err", err.max())
    assert err.max() < 2e-7


@pytest.mark.parametrize("prec,bound", [("bf16", 3e-2), ("fp16", 4e-3)])
def test_pair_16bit_error(xavier_w, prec, bound):
    x = workloads.uniform_cube(100, 0.05, 0)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    r = _disp(4096, 3)
    got = m.eval_pair(r, prec).astype(np.float64)
    ref = O.mlp_f64(xavier_w, r.astype(np.float64))
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(prec, "pair rel err: median", np.median(err), "max", err.max())
    assert err.max() < bound


def test_lists_bit_exact(model_small):
    a = model_small["m"].export()
    b = model_small["om"].export()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_dot_fp32_parity(model_small, xavier_w):
    m, om, f = model_small["m"], model_small["om"], model_small["f"]
    u = m.dot(f)
    ref = om.dot(xavier_w, f)
    err = O.rel_l2(u, ref)
    print("fp32 dot rel err vs oracle", err)
    assert err <= 1e-5
    np.testing.assert_array_equal(m.far_ranks(), om.far_ranks())


def test_dot_fp32_persistent_far_field_parity(model_small, xavier_w, monkeypatch):
    """The persistent far-field kernel (one workgroup runs a block's whole
    ACA) gives the result of the batched iterations."""
    monkeypatch.setenv("HH_FAR_MODE", "0")
    x, f, om = model_small["x"], model_small["f"], model_small["om"]
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50)
    m.update_coord()
    err = O.rel_l2(m.dot(f), om.dot(xavier_w, f))
    print("persistent fp32 dot rel err vs oracle", err)
    assert err <= 1e-5
    np.testing.assert_array_equal(m.far_ranks(), om.far_ranks())


def test_dot_enqueue_states_and_graph_replays(model_small, xavier_w, monkeypatch):
    """hh_model_dot_capturable / hh_model_dot_enqueue: refused without tree
    reuse, in batched mode under host control and before the persistent kernel is set up for
    the current iteration cap; otherwise the enqueued matvec, directly and
    as a HIP graph replayed with eager work in between, equals the eager
    one.  The signature follows the ACA settings."""
    import torch
    x, f, om = model_small["x"], model_small["f"], model_small["om"]
    dev = torch.device("cuda:0")
    f_dev = torch.from_numpy(f).to(dev)
    u_dev = torch.zeros_like(f_dev)

    def make():
        m = abi.Model(x, 64)
        m.load_weights(xavier_w)
        m.set(epsilon=1e-5, max_iter=50)
        m.update_coord()
        return m

    monkeypatch.setenv("HH_FAR_MODE", "1")
    monkeypatch.setenv("HH_ACA_CONTROL", "host")  # batched, stop flags read back per iteration
    mb = make()
    mb.set(tree_reuse=True)
    mb.dot(f)
    assert not mb.dot_capturable()[0]
    monkeypatch.delenv("HH_ACA_CONTROL")

    monkeypatch.setenv("HH_FAR_MODE", "0")
    m = make()
    m.dot(f)
    assert not m.dot_capturable()[0]  # tree reuse off: update_coord rebuilds on the host
    with pytest.raises(RuntimeError, match="capturable"):
        m.dot_enqueue(u_dev.data_ptr(), f_dev.data_ptr())
    m.set(tree_reuse=True)
    want = m.dot(f)
    assert O.rel_l2(want, om.dot(xavier_w, f)) <= 1e-5
    cap, sig = m.dot_capturable()
    assert cap
    m.dot_enqueue(u_dev.data_ptr(), f_dev.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert O.rel_l2(u_dev.cpu().numpy(), want) <= 1e-6  # FP64 atomics: order only

    g = torch.cuda.CUDAGraph()
    u_g = torch.zeros_like(f_dev)
    with torch.cuda.graph(g):
        m.dot_enqueue(u_g.data_ptr(), f_dev.data_ptr(), torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        u_g.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        err = O.rel_l2(u_g.cpu().numpy(), want)
        print(f"graph replay {rep}: rel diff {err:.2e}")
        assert err <= 1e-6
        assert O.rel_l2(m.dot(f), want) <= 1e-6  # eager matvec between replays

    m.set(epsilon=1e-4)
    cap2, sig2 = m.dot_capturable()
    assert cap2 and sig2 != sig  # the tolerance is a launch argument
    m.set(max_iter=30)
    assert not m.dot_capturable()[0]  # per-block pools sized for 50 iterations
    m.dot(f)
    assert m.dot_capturable()[0]


def test_tune_far_field(model_small, xavier_w):
    """hh_model_tune_far_field times every candidate (the persistent kernel
    only with at most 4 far blocks per compute unit, 256 on MI355X; batched
    with chunks of 64, 128, 256 items) and keeps the fastest; the tuned
    model's matvec still meets the FP32 bar against the oracle."""
    import torch
    x, f, om = model_small["x"], model_small["f"], model_small["om"]
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50)
    m.update_coord()
    f_dev = torch.from_numpy(f).to(torch.device("cuda:0"))
    rows = m.tune_far_field(f_dev.data_ptr(), torch.cuda.current_stream().cuda_stream)
    print("tuning rows (mode, chunk, far ms):", rows)
    nfar = len(abi.tree_build(x, 64)["far"])
    assert sorted(r[1] for r in rows if r[0] == 1) == [64, 128, 256]
    assert any(r[0] == 0 for r in rows) == (nfar <= 4 * 256)
    assert all(r[2] > 0 for r in rows)
    assert O.rel_l2(m.dot(f), om.dot(xavier_w, f)) <= 1e-5


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_near_shapes_and_tuning(model_small, xavier_w, prec):
    """Every tile shape of the 16-bit near kernel (16-pair tiles per MLP call
    x threads per workgroup) and every workgroup count computes the same
    matvec up to the FP64 summation order; hh_model_tune_near times the
    shapes on all compute units and, with a far field beside the near
    field, shapes 0..2 on 5/8, 1/2 and 3/8 of them, and keeps the fastest; the
    FP32 mode has nothing to tune."""
    x, f = model_small["x"], model_small["f"]
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50, precision=prec)
    m.update_coord()
    us = []
    for shape in range(4):
        m.set_near_shape(shape)
        us.append(m.dot(f))
    for shape in range(1, 4):
        assert O.rel_l2(us[shape], us[0]) <= 1e-12, shape
    with pytest.raises(RuntimeError):
        m.set_near_shape(4)
    m.set_near_shape(0)
    for wg in (1, 3, 100000):  # one workgroup drains the whole task queue
        m.set_near_workgroups(wg)
        assert O.rel_l2(m.dot(f), us[0]) <= 1e-12, wg
    m.set_near_workgroups(0)
    with pytest.raises(RuntimeError):
        m.set_near_workgroups(-1)
    fd = abi.DeviceArray.of(f)
    rows = m.tune_near(fd.ptr)
    print(f"near launches ({prec}):", rows)
    assert m.far_ranks().size > 0  # the capped candidates are timed
    assert [r[0] for r in rows] == [0, 1, 2, 3] + [0, 1, 2] * 3
    assert [(r[1], r[2]) for r in rows[:4]] == [(4, 512), (2, 512), (2, 1024), (4, 256)]
    cus = rows[0][3]
    assert [r[3] for r in rows] == [cus] * 4 + [cus * 5 // 8] * 3 + [cus // 2] * 3 + [cus * 3 // 8] * 3
    assert all(r[4] > 0 for r in rows)
    assert O.rel_l2(m.dot(f), us[0]) <= 1e-12
    m.set(precision="fp32")
    assert m.tune_near(fd.ptr) == []


def test_dense_dot_parity(model_small, xavier_w):
    m, om, f = model_small["m"], model_small["om"], model_small["f"]
    err = O.rel_l2(m.dense_dot(f), om.dense_dot(xavier_w, f))
    print("dense dot rel err", err)
    assert err <= 1e-5


def test_check_fp64_matches_oracle(model_small, xavier_w):
    m, om, f, x = model_small["m"], model_small["om"], model_small["f"], model_small["x"]
    tg = np.arange(0, x.shape[0], 37)
    e0 = O.rel_l2(m.check_fp64(f, tg, 0), om.dense64_blocks(xavier_w, f, tg))
    e1 = O.rel_l2(m.check_fp64(f, tg, 1), O.dense64_ordered(xavier_w, x, f, tg))
    print("fp64 check vs oracle", e0, e1)
    assert e0 < 1e-10 and e1 < 1e-10


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_chosen_form_accuracy(model_small, prec):
    m, f, x = model_small["m"], model_small["f"], model_small["x"]
    tg = np.arange(0, x.shape[0], 5)
    ref = m.check_fp64(f, tg, 0)
    m.set(precision=prec)
    try:
        u = m.dot(f)
    finally:
        m.set(precision="fp32")
    u32 = m.dot(f)
    e = O.rel_l2(u[tg], ref)
    e32 = O.rel_l2(u32[tg], ref)
    print(prec, "error vs fp64 blocks", e, "fp32 path", e32)
    assert e <= 1e-3



@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_chosen_form_persistent_far_field(model_small, xavier_w, monkeypatch, prec):
    """The 16-bit persistent far-field kernel (one workgroup per block, FP32
    factor slots) against the FP64 evaluation and against the batched
    iterations (BF16 factor pools): both within the 1e-3 bar, and close to
    each other."""
    x, f = model_small["x"], model_small["f"]
    tg = np.arange(0, x.shape[0], 5)
    ref = model_small["m"].check_fp64(f, tg, 0)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HH_FAR_MODE", mode)
        m = abi.Model(x, 64)
        m.load_weights(xavier_w)
        m.set(epsilon=1e-5, max_iter=50, precision=prec)
        m.update_coord()
        out[mode] = m.dot(f)
        assert np.isfinite(out[mode]).all()
        e = O.rel_l2(out[mode][tg], ref)
        print(f"{prec} far mode {mode}: error vs fp64 blocks {e:.2e}")
        assert e <= 1e-3
    d = O.rel_l2(out["0"], out["1"])
    print(f"{prec}: persistent vs batched {d:.2e}")
    assert d <= 1e-3


@pytest.mark.parametrize("chunk", [16, 33])
def test_dot_fp32_parity_multichunk_blocks(model_small, xavier_w, monkeypatch, chunk):
    """ACA steps and apply cut far blocks into chunks of HH_ACA_CHUNK items;
    with small chunks every large far block spans several chunks and the
    per-chunk partials are combined across workgroups."""
    monkeypatch.setenv("HH_ACA_CHUNK", str(chunk))
    x, f, om = model_small["x"], model_small["f"], model_small["om"]
    exp = om.export()
    sz = exp["tree"][:, 3] - exp["tree"][:, 2]
    assert sz[exp["far"]].max() > 2 * chunk
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50)
    m.update_coord()
    err = O.rel_l2(m.dot(f), om.dot(xavier_w, f))
    print(f"chunk {chunk}: fp32 dot rel err vs oracle", err)
    assert err <= 1e-5
    np.testing.assert_array_equal(m.far_ranks(), om.far_ranks())
    m.set(precision="fp16")
    tg = np.arange(0, x.shape[0], 5)
    e = O.rel_l2(m.dot(f)[tg], m.check_fp64(f, tg, 0))
    print(f"chunk {chunk}: fp16 error vs fp64 blocks", e)
    assert e <= 1e-3


@pytest.mark.parametrize("eps,chunk,mode", [(0.05, None, "1"), (1e-2, None, "1"), (1e-3, None, "1"),
                                            (1e-2, 16, "1"), (1e-2, None, "0")])
def test_dot_fp32_parity_tolerance_stops(model_small, xavier_w, monkeypatch, eps, chunk, mode):
    """Blocks that stop on the tolerance test (nu2 < mu2 eps^2) rather than on
    the iteration count: the batched path decides that test one iteration
    late, from the cross terms taken in the next residual pass; ranks and
    velocities must still be the oracle's.  mode "0" is the persistent
    kernel.

    The FP32 pair function differs from the oracle's by ~1e-6 per entry
    (summation order).  A late ACA residual is a difference of entries
    ~1/eps larger than itself, so at eps = 1e-3 that difference reaches
    ~1e-3 of the residual, comparable to the closest stop margins and pivot
    gaps of this case (2.5e-3 and 7e-5, from an FP64 restatement); there a
    rank may differ in a few blocks and only the velocities are held to
    the 1e-5 bar.  At eps >= 1e-2 every rank must match."""
    if chunk is not None:
        monkeypatch.setenv("HH_ACA_CHUNK", str(chunk))
    monkeypatch.setenv("HH_FAR_MODE", mode)
    x, f = model_small["x"], model_small["f"]
    om = O.OracleModel(x, 64, epsilon=eps, max_iter=50)
    om.update_coord()
    ref = om.dot(xavier_w, f)
    exp = om.export()
    sz = exp["tree"][:, 3] - exp["tree"][:, 2]
    hard = np.minimum(50, np.minimum(sz[exp["far"][:, 0]], sz[exp["far"][:, 1]]))
    assert (om.far_ranks() < hard).sum() > 0  # the case exercises tolerance stops
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=eps, max_iter=50)
    m.update_coord()
    err = O.rel_l2(m.dot(f), ref)
    ranks, oranks = m.far_ranks(), om.far_ranks()
    same = float(np.mean(ranks == oranks))
    print(f"eps {eps} chunk {chunk} mode {mode}: fp32 dot rel err vs oracle {err}, "
          f"ranks equal {same:.4f}")
    assert err <= 1e-5
    if eps >= 1e-2:
        np.testing.assert_array_equal(ranks, oranks)
    else:
        assert same >= 0.97 and np.abs(ranks - oranks).max() <= 8


def test_tree_reuse_keeps_lists_and_tracks_positions(xavier_w):
    """Opt-in tree reuse: after a first build, update_coord keeps the lists
    and refreshes the positions.  The matvec at the moved positions then
    equals the FP64 evaluation of the kept block structure at those
    positions, to the ACA tolerance, and moving back restores the first
    result (up to the order of the FP64 atomic sums)."""
    x0 = workloads.uniform_cube(1200, 0.05, 0)
    g = np.random.default_rng(11)
    x1 = (x0 + 0.05 * g.standard_normal(x0.shape)).astype(np.float32)
    f = workloads.gravity(x0.shape[0])
    m = abi.Model(x0, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50, tree_reuse=True)
    m.update_coord()
    e0 = m.export()
    u0 = m.dot(f)
    m.update_coord(x1)
    e1 = m.export()
    for k in e0:
        if k != "aux":
            np.testing.assert_array_equal(e0[k], e1[k], err_msg=k)
    u1 = m.dot(f)
    tg = np.arange(0, x0.shape[0], 7)
    err = O.rel_l2(u1[tg], m.check_fp64(f, tg, 0))
    moved = O.rel_l2(u1, u0)
    print("tree reuse: error vs fp64 of the kept blocks", err, "change from x0", moved)
    assert err <= 1e-3 and moved > 0.0
    m.update_coord(x0)
    np.testing.assert_allclose(m.dot(f), u0, rtol=1e-5, atol=1e-6 * np.abs(u0).max())


def test_far_mode_auto(model_small, xavier_w, monkeypatch):
    # unset: the persistent kernel when far blocks <= compute units (1200
    # particles: a few hundred blocks at most), same result as both forced modes
    x, f, om = model_small["x"], model_small["f"], model_small["om"]
    ref = om.dot(xavier_w, f)
    out = {}
    for mode in (None, "0", "1"):
        if mode is None:
            monkeypatch.delenv("HH_FAR_MODE", raising=False)
        else:
            monkeypatch.setenv("HH_FAR_MODE", mode)
        m = abi.Model(x, 64)
        m.load_weights(xavier_w)
        m.set(epsilon=1e-5, max_iter=50)
        m.update_coord()
        out[mode] = m.dot(f)
        assert O.rel_l2(out[mode], ref) <= 1e-5, mode
    nfar = len(abi.tree_build(x, 64)["far"])  # 256 compute units on MI355X
    expect = "0" if nfar <= 256 else "1"
    print("far blocks:", nfar, "-> mode", expect)
    assert O.rel_l2(out[None], out[expect]) <= 1e-7


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("case", ["one", "two_coincident", "three_within_cutoff"])
def test_dot_tiny_systems_closed_form(xavier_w, prec, case):
    # Pairs closer than 1e-5 interact through the identity block
    # (python/convert.py:14-19), the self pair included, so every particle
    # moves with the sum of the forces on the group.
    x = {"one": [[0.5, -1.0, 2.0]],
         "two_coincident": [[1.0, 2.0, 3.0]] * 2,
         "three_within_cutoff": [[0.0, 0.0, 0.0], [3e-6, 0.0, 0.0], [0.0, -4e-6, 2e-6]]}[case]
    x = np.asarray(x, np.float32)
    f = np.random.default_rng(len(x)).normal(size=x.shape).astype(np.float32)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50, precision=prec)
    m.update_coord()
    u = m.dot(f)
    np.testing.assert_allclose(u, np.broadcast_to(f.sum(0), x.shape), rtol=1e-6, atol=1e-7)
    if prec == "fp32":
        np.testing.assert_allclose(m.dense_dot(f), u, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("n", [2, 50, 65, 130])
def test_dot_fp32_parity_below_and_near_one_leaf(xavier_w, n):
    # a root that stays a leaf (n <= 64), one split (65) and two levels (130)
    x = workloads.uniform_cube(n, 0.05, n)
    f = np.random.default_rng(n).normal(size=x.shape).astype(np.float32)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-5, max_iter=50)
    m.update_coord()
    om = O.OracleModel(x, 64, epsilon=1e-5, max_iter=50)
    om.update_coord()
    assert O.rel_l2(m.dot(f), om.dot(xavier_w, f)) <= 1e-5
    assert O.rel_l2(m.dense_dot(f), om.dense_dot(xavier_w, f)) <= 1e-5


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_singular_pivot_blocks_stay_finite(xavier_w, prec, monkeypatch):
    """A pair function that vanishes away from r = 0 (W3 = 0, b3 = -I, so
    y + I = 0) makes every far block's pivot block exactly singular.  Both
    far-field paths give the singular pivot a zero inverse: no inf/NaN
    reaches u, which is then the self-interaction alone (u_i = f_i), and the
    persistent kernel stops every block at the same rank as the batched
    iterations."""
    p = weights.unpack(xavier_w)
    p = {k: v.copy() for k, v in p.items()}
    p["W3"][:] = 0.0
    p["b3"][:] = [-1, 0, 0, 0, -1, 0, 0, 0, -1]
    w = weights.pack(p)
    x = workloads.uniform_cube(400, 0.05, 9)
    f = np.random.default_rng(9).normal(size=x.shape).astype(np.float32)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("HH_FAR_MODE", mode)
        m = abi.Model(x, 32)
        m.load_weights(w)
        m.set(epsilon=1e-4, max_iter=12, precision=prec)
        m.update_coord()
        u = m.dot(f)
        assert len(m.far_ranks()) > 0
        assert np.isfinite(u).all(), mode
        np.testing.assert_allclose(u, f, rtol=1e-6, atol=1e-6)
        out[mode] = m.far_ranks()
    np.testing.assert_array_equal(out["0"], out["1"])
