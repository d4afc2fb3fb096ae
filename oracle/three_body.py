"""numpy restatement of the neighbour lists and the three-body / self
corrections (TEST INFRASTRUCTURE).

Only tests/ import this module, as the checker of the device lists and MLP
kernels.

- ``neighbour_lists``: NeighborLists::BuildTargetNeighborLists
  (include/NeighborLists.hpp:166-216) with the kd-tree's radius test
  (oracle/forcing.neighbours), sources = the particles or, with a periodic
  domain, the images of BuildSourceSites (:49-164) whose margin is the
  two-body radius (:61); neighbours mapped to targets, the first entry equal
  to the target dropped.  Each list is returned sorted.
- ``two_body_edges``: BuildTwoBodyInfo (:466-510), including its second
  lookup through the image-to-target map (:497-500).
- ``three_body_edges``: BuildThreeBodyInfo (:512-588): (i, n, p) for n in
  N(i), p in N(n), p != i; self edges (i, n) for n in N(i).
- ``net``: Linear(d,256) Tanh Linear(256,1024) Tanh Linear(1024,512) Tanh
  Linear(512,9) in float64 (python/HIGNN/model_structure.py:41-68).
- ``velocity``: the three-body and self parts of HIGNN_mdoel.forward
  (python/HIGNN/model_structure.py:74-140) with empty two-body edges, as
  velocity_update_without_2body calls it (python/gravity_field.py:34-59).

Pinned by tests/test_three_body_oracle.py: the net against torch's
nn.Sequential in float64, list sizes of a cubic lattice against their closed
form, the triplets against a set construction from scipy's cKDTree, and the
velocity's translation invariance and linearity in the forces.
"""
import numpy as np

from oracle import forcing as F

f32 = np.float32


def neighbour_lists(x, eps, periodic=False, lo=None, hi=None, margin=0.0):
    x = np.asarray(x, dtype=f32)
    if periodic:
        src, tgt = F.source_images(x, margin, lo, hi)
    else:
        src, tgt = x, np.arange(x.shape[0])
    nb = F.neighbours(x, src, eps)
    out = []
    for i in range(x.shape[0]):
        t = tgt[nb[i]]
        hit = np.nonzero(t == i)[0]
        if hit.size:
            t = np.delete(t, hit[0])
        out.append(np.sort(t))
    return out, tgt


def two_body_edges(x, eps, periodic=False, lo=None, hi=None):
    lists, tgt = neighbour_lists(x, eps, periodic, lo, hi, margin=eps)
    rows = [(i, int(tgt[j])) for i, l in enumerate(lists) for j in l]
    return np.array(rows, dtype=np.int64).reshape(-1, 2).T


def three_body_edges(lists):
    """(3,E) triplets and (2,E) self edges, grouped by target, ascending."""
    trip, self_e = [], []
    for i, l in enumerate(lists):
        for n in l:
            for p in lists[n]:
                if p != i:
                    trip.append((i, n, p))
            self_e.append((i, n))
    return (np.array(trip, dtype=np.int64).reshape(-1, 3).T,
            np.array(self_e, dtype=np.int64).reshape(-1, 2).T)


def net(w, din, xin):
    """The net in float64 for inputs xin (m, din); w is the flat weight vector."""
    w = np.asarray(w, dtype=np.float64)
    shapes = [(256, din), (256,), (1024, 256), (1024,), (512, 1024), (512,), (9, 512), (9,)]
    parts, off = [], 0
    for s in shapes:
        k = int(np.prod(s))
        parts.append(w[off:off + k].reshape(s))
        off += k
    h = np.asarray(xin, dtype=np.float64)
    for layer in range(4):
        W, b = parts[2 * layer], parts[2 * layer + 1]
        h = h @ W.T + b
        if layer < 3:
            h = np.tanh(h)
    return h


def velocity(x, f, w3, wself, eps, chunk=20000, targets=None):
    """v_3body + v_self (float64, (n,3)) on the lists within eps; with
    ``targets``, only the edges into those particles are evaluated (the other
    rows stay zero)."""
    x = np.asarray(x, dtype=f32)
    f = np.asarray(f, dtype=f32)
    lists, _ = neighbour_lists(x, eps)
    trip, self_e = three_body_edges(lists)
    if targets is not None:
        keep = np.zeros(x.shape[0], dtype=bool)
        keep[np.asarray(targets)] = True
        trip = trip[:, keep[trip[2]]]
        self_e = self_e[:, keep[self_e[1]]]
    v = np.zeros((x.shape[0], 3))
    X = x.astype(np.float64)
    F64 = f.astype(np.float64)
    for a in range(0, trip.shape[1], chunk):
        i, n, p = trip[:, a:a + chunk]
        d1 = X[n] - X[i]
        d2 = X[p] - X[n]
        y = net(w3, 6, np.concatenate([d1, d2], axis=1))
        y /= np.linalg.norm(d1, axis=1, keepdims=True)
        y /= np.linalg.norm(d2, axis=1, keepdims=True)
        m = y.reshape(-1, 3, 3)
        np.add.at(v, p, np.einsum("eab,eb->ea", m, F64[i]))
    for a in range(0, self_e.shape[1], chunk):
        i, n = self_e[:, a:a + chunk]
        d = X[i] - X[n]
        y = net(wself, 3, d)
        r = np.linalg.norm(d, axis=1, keepdims=True)
        m = (y / r / r).reshape(-1, 3, 3)
        np.add.at(v, n, np.einsum("eab,eb->ea", m, F64[n]))
    return v
