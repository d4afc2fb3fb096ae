"""Pins of oracle/three_body.py (CPU only)."""
import numpy as np
import pytest

from hignn_amd import weights as W
from oracle import three_body as T

f32 = np.float32


def lattice(m, dx=3.0, jitter=0.0, seed=0):
    g = np.arange(m) * dx
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    if jitter:
        x = x + np.random.default_rng(seed).uniform(-jitter, jitter, x.shape)
    return x.astype(f32)


def test_net_matches_torch_sequential():
    torch = pytest.importorskip("torch")
    for din in (6, 3):
        w = W.xavier(3, W.net_shapes(din))
        parts = W.unpack(w, W.net_shapes(din))
        seq = torch.nn.Sequential(torch.nn.Linear(din, 256), torch.nn.Tanh(),
                                  torch.nn.Linear(256, 1024), torch.nn.Tanh(),
                                  torch.nn.Linear(1024, 512), torch.nn.Tanh(),
                                  torch.nn.Linear(512, 9)).double()
        with torch.no_grad():
            for k, lin in enumerate([seq[0], seq[2], seq[4], seq[6]]):
                lin.weight.copy_(torch.from_numpy(parts[f"W{k + 1}"].astype(np.float64)))
                lin.bias.copy_(torch.from_numpy(parts[f"b{k + 1}"].astype(np.float64)))
            xin = np.random.default_rng(din).normal(size=(64, din)) * 4
            want = seq(torch.from_numpy(xin)).numpy()
        np.testing.assert_allclose(T.net(w, din, xin), want, rtol=1e-12, atol=1e-12)


def test_lattice_counts_closed_form():
    # spacing 3, radius 5: the 6 face (3) and 12 edge (4.24) neighbours, not the corners (5.2)
    x = lattice(7)
    lists, _ = T.neighbour_lists(x, 5.0)
    idx = lambda a, b, c: (a * 7 + b) * 7 + c
    centre = idx(3, 3, 3)
    assert len(lists[centre]) == 18
    assert len(lists[idx(0, 0, 0)]) == 6  # corner: 3 face + 3 edge
    trip, self_e = T.three_body_edges(lists)
    # every neighbour of the centre has 18 neighbours itself, one of them the centre
    assert (trip[0] == centre).sum() == 18 * 17
    assert self_e.shape[1] == sum(len(l) for l in lists)
    assert (trip[0] != trip[2]).all() and (trip[0] != trip[1]).all() and (trip[1] != trip[2]).all()


def test_triplets_against_kdtree_sets():
    from scipy.spatial import cKDTree
    x = np.random.default_rng(1).uniform(0, 12, size=(150, 3)).astype(f32)
    lists, _ = T.neighbour_lists(x, 2.5)
    trip, self_e = T.three_body_edges(lists)
    tree = cKDTree(x.astype(np.float64))
    nb = [set(tree.query_ball_point(x[i].astype(np.float64), 2.5)) - {i} for i in range(150)]
    want = {(i, n, p) for i in range(150) for n in nb[i] for p in nb[n] if p != i}
    assert set(map(tuple, trip.T.tolist())) == want
    assert len(want) == trip.shape[1]
    assert set(map(tuple, self_e.T.tolist())) == {(i, n) for i in range(150) for n in nb[i]}


def test_two_body_edges_non_periodic_are_the_neighbour_pairs():
    x = np.random.default_rng(2).uniform(0, 10, size=(120, 3)).astype(f32)
    e = T.two_body_edges(x, 2.0)
    lists, _ = T.neighbour_lists(x, 2.0)
    assert e.shape[1] == sum(len(l) for l in lists)
    pairs = set(map(tuple, e.T.tolist()))
    assert all((j, i) in pairs for i, j in pairs)  # symmetric without images


def test_velocity_invariances():
    x = lattice(4, jitter=0.3)
    x = (np.round(x * 1024) / 1024).astype(f32)  # on a 2^-10 grid: x + 64 stays exact
    rng = np.random.default_rng(3)
    f = (np.round(rng.normal(size=x.shape) * 64) / 64).astype(f32)  # f + g exact in float32
    w3, ws = W.xavier(1, W.net_shapes(6)), W.xavier(2, W.net_shapes(3))
    v = T.velocity(x, f, w3, ws, 5.0)
    assert np.abs(v).max() > 0
    np.testing.assert_array_equal(T.velocity(x, 2 * f, w3, ws, 5.0), 2 * v)
    g = (np.round(rng.normal(size=x.shape) * 64) / 64).astype(f32)
    np.testing.assert_allclose(T.velocity(x, f + g, w3, ws, 5.0),
                               v + T.velocity(x, g, w3, ws, 5.0), atol=1e-9 * np.abs(v).max())
    # the shifted coordinates and all their differences are exact in float32
    np.testing.assert_allclose(T.velocity(x + f32(64.0), f, w3, ws, 5.0), v, rtol=1e-12, atol=1e-12)
