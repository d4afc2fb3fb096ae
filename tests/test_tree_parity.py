"""The product's host tree and block lists (libhh hh_tree_build, called
through the C ABI) against the oracle: bit-exact rows, permutation, boxes,
close/far lists and leaf order."""
import numpy as np
import pytest

from hignn_amd import abi, workloads
from oracle import oracle as O


def lattice():
    g = np.arange(10.0) * 2.5
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


CASES = {
    "uniform_2k_leaf64": (lambda: workloads.uniform_cube(2000, 0.05, 0), 64, True),
    "uniform_2k_nosym": (lambda: workloads.uniform_cube(2000, 0.05, 0), 64, False),
    "blobs_20k_leaf256": (lambda: workloads.gaussian_blobs(20000, 4, 15.0, 0.1, 2), 256, True),
    "filaments_2k": (lambda: workloads.filaments(40, 50, 2.2, 0.08, 3), 64, True),
    "lattice_ties": (lattice, 64, True),
    "n_below_leaf": (lambda: workloads.uniform_cube(50, 0.05, 1), 64, True),
    "n_equals_leaf_plus_one": (lambda: workloads.uniform_cube(65, 0.05, 1), 64, True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_tree_lists_bit_exact(name):
    make, leaf, sym = CASES[name]
    x = make()
    a = abi.tree_build(x, leaf, use_symmetry=sym)
    om = O.OracleModel(x, leaf, use_symmetry=sym)
    om.update_coord()
    b = om.export()
    for k in ("tree", "perm", "aux", "close", "far", "leaves"):
        assert a[k].shape == b[k].shape, k
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_max_relative_coord_bounds_far_nodes():
    """maxRelativeCoord (src/CloseFarCheck.cpp:199-203, 240-257) keeps every
    far node smaller than the bound."""
    x = workloads.uniform_cube(6000, 0.1, 5)
    a = abi.tree_build(x, 64, max_relative_coord=300)
    tree = a["tree"].astype(np.int64)
    sz = tree[:, 3] - tree[:, 2]
    assert sz[a["far"]].max() < 300
    om = O.OracleModel(x, 64, max_relative_coord=300)
    om.update_coord()
    np.testing.assert_array_equal(a["far"], om.export()["far"])


def test_max_relative_coord_at_leaf_size_terminates():
    """maxRelativeCoord <= leaf size: leaves are taken as far nodes."""
    x = workloads.uniform_cube(3000, 0.1, 6)
    a = abi.tree_build(x, 64, max_relative_coord=64)
    om = O.OracleModel(x, 64, max_relative_coord=64)
    om.update_coord()
    b = om.export()
    for k in ("close", "far", "leaves"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_invalid_arguments_raise():
    with pytest.raises(RuntimeError):
        abi.tree_build(np.zeros((0, 3), np.float32), 64)
    with pytest.raises(RuntimeError):
        abi.tree_build(np.zeros((10, 3), np.float32), 0)
