"""The cluster tree and close/far block lists built on the device
(hh_tree_build_device, hh_model_set_tree_build) against the oracle and the
host build: rows, permutation, boxes, close/far lists and leaf order bit for
bit, for single systems, forests of independent instances and the model's
update_coord paths (GPU)."""
import time

import numpy as np
import pytest

from hignn_amd import abi, weights, workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KEYS = ("tree", "perm", "aux", "close", "far", "leaves")


def lattice():
    g = np.arange(10.0) * 2.5
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def signed_zero_lattice():
    x = lattice() - 11.25
    x[::7, 0] = -0.0
    x[1::7, 0] = 0.0
    return x


CASES = {
    "uniform_2k_leaf64": (lambda: workloads.uniform_cube(2000, 0.05, 0), 64, True, 500000),
    "uniform_2k_nosym": (lambda: workloads.uniform_cube(2000, 0.05, 0), 64, False, 500000),
    "blobs_20k_leaf256": (lambda: workloads.gaussian_blobs(20000, 4, 15.0, 0.1, 2), 256, True, 500000),
    "c3_blobs_20k_R1000": (lambda: workloads.positions("c3_blobs_1m", n=20000), 256, True, 1000),
    "filaments_2k": (lambda: workloads.filaments(40, 50, 2.2, 0.08, 3), 64, True, 500000),
    "lattice_ties": (lattice, 64, True, 500000),
    "signed_zeros": (signed_zero_lattice, 32, True, 500000),
    "n_below_leaf": (lambda: workloads.uniform_cube(50, 0.05, 1), 64, True, 500000),
    "n_equals_leaf_plus_one": (lambda: workloads.uniform_cube(65, 0.05, 1), 64, True, 500000),
    "max_rel_coord_300": (lambda: workloads.uniform_cube(6000, 0.1, 5), 64, True, 300),
    "max_rel_coord_at_leaf": (lambda: workloads.uniform_cube(3000, 0.1, 6), 64, True, 64),
}


def assert_same(a, b):
    for k in KEYS:
        assert a[k].shape == b[k].shape, k
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_tree_bit_exact(name):
    make, leaf, sym, R = CASES[name]
    x = make()
    print(f"device tree: case {name} n={x.shape[0]}", flush=True)
    d = abi.tree_build_device(x, leaf, use_symmetry=sym, max_relative_coord=R)
    om = O.OracleModel(x, leaf, use_symmetry=sym, max_relative_coord=R)
    om.update_coord()
    assert_same(d, om.export())
    assert_same(d, abi.tree_build(x, leaf, use_symmetry=sym, max_relative_coord=R))


def test_device_forest_equals_merged_host_build():
    """Independent instances in one device build (one set of level passes)
    against the model's host build of the same instances."""
    x, ns = workloads.instances(count=7, nmin=40, nmax=3000, seed=11)
    print(f"device forest: {len(ns)} instances, n={x.shape[0]}", flush=True)
    d = abi.tree_build_device(x, 64, counts=ns)
    m = abi.Model(x, 64, instances=ns)
    m.load_weights(weights.smooth(0))
    m.set(tree_build="host")
    m.update_coord()
    assert_same(d, m.export())


@pytest.mark.parametrize("inst", [False, True])
def test_model_device_build_same_lists_and_velocity(inst):
    """update_coord with the device build: the same tree and lists as the host
    build, hence the same matvec."""
    if inst:
        x, ns = workloads.instances(count=4, nmin=500, nmax=4000, seed=5)
    else:
        x, ns = workloads.uniform_cube(5000, 0.1, 3), None
    f = workloads.gravity(x.shape[0])
    w = weights.smooth(0)
    out = {}
    for mode in ("host", "device"):
        m = abi.Model(x, 64, instances=ns)
        m.load_weights(w)
        m.set(epsilon=1e-4, max_iter=30, tree_build=mode)
        m.update_coord()
        out[mode] = (m.export(), m.dot(f), m.far_ranks())
    assert_same(out["device"][0], out["host"][0])
    np.testing.assert_array_equal(out["device"][2], out["host"][2])
    err = np.linalg.norm(out["device"][1] - out["host"][1]) / np.linalg.norm(out["host"][1])
    assert err <= 1e-6  # same blocks and ranks: atomic summation order only


def test_update_coord_device_builds_from_device_positions():
    """update_coord_device with the device build reads the caller's device
    array (no host copy) and tracks moved positions; a later host-path
    rebuild sees the same positions."""
    x0 = workloads.uniform_cube(4000, 0.1, 9)
    rng = np.random.default_rng(1)
    x1 = (x0 + rng.normal(scale=0.5, size=x0.shape)).astype(np.float32)
    w = weights.smooth(0)
    m = abi.Model(x0, 64)
    m.load_weights(w)
    m.set(epsilon=1e-4, max_iter=30, tree_build="device")
    m.update_coord()
    xd = abi.DeviceArray.of(x1)
    m.update_coord_device(xd.ptr)
    abi.synchronize()
    assert_same(m.export(), abi.tree_build(x1, 64))
    ref = abi.Model(x1, 64)
    ref.load_weights(w)
    ref.set(epsilon=1e-4, max_iter=30, tree_build="host")
    ref.update_coord()
    f = workloads.gravity(x1.shape[0])
    u, ur = m.dot(f), ref.dot(f)
    assert np.linalg.norm(u - ur) / np.linalg.norm(ur) <= 1e-6
    # switching the symmetry flag makes the next dot rebuild on the host from
    # the host copy, which must hold x1 (refreshed from the device positions)
    m.set(use_symmetry=False, tree_build="host")
    m.dot(f)
    assert_same(m.export(), abi.tree_build(x1, 64, use_symmetry=False))


def test_device_build_at_1m_blobs_properties():
    """Full-size C3 positions (10^6 particles in 200 blobs, leaf 128): the
    device build equals the host build; both timed."""
    x = workloads.positions("c3_blobs_1m")
    t0 = time.perf_counter()
    h = abi.tree_build(x, 128)
    th = time.perf_counter() - t0
    abi.tree_build_device(x, 128)  # warm-up (allocations)
    t0 = time.perf_counter()
    d = abi.tree_build_device(x, 128)
    td = time.perf_counter() - t0
    print(f"tree+lists at N=1e6: host {th:.3f} s, device {td:.3f} s "
          f"({len(d['close'])} close, {len(d['far'])} far pairs)", flush=True)
    assert_same(d, h)
