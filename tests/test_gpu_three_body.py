"""Neighbour lists and three-body / self corrections on the device against
oracle/three_body.py (GPU)."""
import numpy as np
import pytest

from hignn_amd import weights as W
from oracle import three_body as T

pytestmark = pytest.mark.gpu

f32 = np.float32


def hignn():
    from hignn_amd import hignn as h
    h.Init()
    return h


def lattice(m, dx=3.0, jitter=0.25, seed=0):
    g = np.arange(m) * dx
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    x = x + np.random.default_rng(seed).uniform(-jitter, jitter, x.shape)
    return x.astype(f32)


def test_two_body_edges_match_oracle():
    h = hignn()
    x = np.random.default_rng(0).uniform(0, 30, size=(2000, 3)).astype(f32)
    nl = h.NeighborLists()
    nl.set_two_body_epsilon(2.5)
    nl.update_coord(x)
    nl.build_two_body_info()
    got = nl.get_two_body_edge_info()
    assert got.dtype == np.uint64
    np.testing.assert_array_equal(got, T.two_body_edges(x, 2.5))
    # by index: every other particle, paired with the given one
    bi = nl.get_two_body_edge_info_by_index(np.array([3, 7], dtype=np.uint64))
    assert bi.shape == (2, 2 * 1999)
    assert (bi[1, :1999] == 3).all() and 3 not in bi[0, :1999]


def test_two_body_edges_periodic_match_oracle():
    h = hignn()
    lo, hi = np.array([0, 0, 0], f32), np.array([20, 20, 20], f32)
    x = np.random.default_rng(1).uniform(lo, hi, size=(800, 3)).astype(f32)
    nl = h.NeighborLists()
    nl.set_two_body_epsilon(2.0)
    nl.set_domain(np.stack([lo, hi]))
    nl.update_coord(x)
    nl.build_two_body_info()
    np.testing.assert_array_equal(nl.get_two_body_edge_info(),
                                  T.two_body_edges(x, 2.0, periodic=True, lo=lo, hi=hi))


def test_three_body_edges_match_oracle():
    h = hignn()
    x = lattice(8)
    nl = h.NeighborLists()
    nl.set_three_body_epsilon(5.0)
    nl.update_coord(x)
    nl.build_three_body_info()
    lists, _ = T.neighbour_lists(x, 5.0)
    trip, self_e = T.three_body_edges(lists)
    np.testing.assert_array_equal(nl.get_three_body_edge_info(), trip)
    np.testing.assert_array_equal(nl.get_three_body_edge_self_info(), self_e)
    # by index: rows reversed, the given targets' edges in the given order
    idx = np.array([100, 5], dtype=np.uint64)
    bi = nl.get_three_body_edge_info_by_index(idx)
    want = np.concatenate([trip[::-1, trip[0] == 100], trip[::-1, trip[0] == 5]], axis=1)
    np.testing.assert_array_equal(bi, want)
    bs = nl.get_three_body_edge_self_info_by_index(idx)
    want = np.concatenate([self_e[::-1, self_e[0] == 100], self_e[::-1, self_e[0] == 5]], axis=1)
    np.testing.assert_array_equal(bs, want)


def model(h, prec):
    tb = h.ThreeBodyModel()
    tb.load_weights(W.xavier(1, W.net_shapes(6)), W.xavier(2, W.net_shapes(3)))
    tb.set_three_body_epsilon(5.0)
    tb.set_precision(prec)
    return tb


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case():
    x = lattice(6)
    f = np.random.default_rng(4).normal(size=x.shape).astype(f32)
    want = T.velocity(x, f, W.xavier(1, W.net_shapes(6)), W.xavier(2, W.net_shapes(3)), 5.0)
    return x, f, want


def test_three_body_velocity_fp32_matches_oracle(case):
    h = hignn()
    x, f, want = case
    tb = model(h, "fp32")
    got = tb.velocity(x, f)
    err = rel_l2(got, want)
    prof = tb.last_profile()
    print(f"fp32: rel L2 {err:.2e}, edges {prof['three_body_edges']} + {prof['self_edges']}")
    assert prof["three_body_edges"] > 30000
    assert err <= 1e-5


def test_three_body_velocity_fp16_error(case):
    h = hignn()
    x, f, want = case
    err = rel_l2(model(h, "fp16").velocity(x, f), want)
    print(f"fp16: rel L2 {err:.2e}")
    assert err <= 5e-3


def test_three_body_batches_and_repeat(case):
    # more edges than one batch (2^18): a 20^3 lattice has ~2.1M triplets
    h = hignn()
    x = lattice(20, seed=5)
    f = np.random.default_rng(6).normal(size=x.shape).astype(f32)
    tb = model(h, "fp16")
    a = tb.velocity(x, f)
    assert tb.last_profile()["three_body_edges"] > (1 << 18)
    b = tb.velocity(x, f)
    assert rel_l2(a, b) < 1e-6  # same lists, FP64 sums in another order at most
    # the oracle on a sample of targets spread over the whole edge list
    sub = np.arange(0, x.shape[0], 397)
    want = T.velocity(x, f, W.xavier(1, W.net_shapes(6)), W.xavier(2, W.net_shapes(3)), 5.0,
                      targets=sub)
    assert rel_l2(a[sub], want[sub]) <= 5e-3
    c = model(h, "fp32").velocity(x, f)
    assert rel_l2(c[sub], want[sub]) <= 1e-5


def test_three_body_edge_cases():
    h = hignn()
    x = lattice(4)
    f = np.ones_like(x)
    tb = h.ThreeBodyModel()
    with pytest.raises(RuntimeError, match="weights"):
        tb.velocity(x, f)
    tb.load_weights(W.xavier(1, W.net_shapes(6)), W.xavier(2, W.net_shapes(3)))
    with pytest.raises(RuntimeError, match="radius"):
        tb.velocity(x, f)
    with pytest.raises(RuntimeError, match="floats"):
        tb.load_weights(W.xavier(1, W.net_shapes(3)), W.xavier(2, W.net_shapes(3)))
    # a radius below the lattice spacing: no edges, zero velocity
    tb.set_three_body_epsilon(1.0)
    v = tb.velocity(x, f)
    assert tb.last_profile()["three_body_edges"] == 0 and tb.last_profile()["self_edges"] == 0
    assert not v.any()
    # a radius that only reaches the face neighbours: self edges but no
    # triplet through a corner (every neighbour of a neighbour except i)
    tb.set_three_body_epsilon(3.5)
    tb.velocity(x, f)
    lists, _ = T.neighbour_lists(x, 3.5)
    trip, self_e = T.three_body_edges(lists)
    prof = tb.last_profile()
    assert prof["three_body_edges"] == trip.shape[1] and prof["self_edges"] == self_e.shape[1]


def test_neighbour_lists_state_errors():
    h = hignn()
    nl = h.NeighborLists()
    with pytest.raises(RuntimeError, match="update_coord"):
        nl.set_two_body_epsilon(1.0)
        nl.build_two_body_info()
    nl.update_coord(lattice(3))
    # before build_three_body_info: empty edge arrays, as the reference's
    # default-constructed views give
    assert nl.get_three_body_edge_info().shape == (3, 0)
    assert nl.get_three_body_edge_self_info().shape == (2, 0)
    with pytest.raises(RuntimeError, match="radius"):
        nl.build_three_body_info()
    nl.set_periodic(True)
    nl.set_three_body_epsilon(2.0)
    with pytest.raises(RuntimeError, match="domain"):
        nl.build_three_body_info()
