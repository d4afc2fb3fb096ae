"""Writes the golden fixtures of the workload parity tests (tests/golden/*.npz)
from the CPU oracle (oracle/hh_oracle.cpp, oracle/integrators.py,
oracle/forcing.py), so that the GPU tests compare against committed data
instead of running the oracle at size.

    python tests/golden/make_fixtures.py [name ...]

Positions are not stored: every configuration is regenerated on the GPU box
from its recipe (hignn_amd.workloads, a pure function of its arguments), and
each fixture stores a checksum of those positions so a changed recipe fails
loudly.  Fixtures:

* c1: N = 2,000, 5% uniform cube, seed 0, leaf 64, eps 1e-5, max_iter 100
  (the reference's default, src/HignnModel.cpp:378; at 50 most blocks stop
  unconverged and ulp-level changes of the pair function move u by ~4e-6),
  Xavier(0) weights, gravity: tree, lists, far ranks and velocities of one
  matvec, and positions after 10 explicit-Euler steps (dt 0.005), each step
  rebuilding the tree (update_coord) before its matvec.
* c2: the C2 recipe (10% uniform cube, seed 1, leaf 256, eps 1e-4, max_iter
  32) at N = 20,000, smooth(0) weights, gravity: lists, ranks, velocities.
* c2_dopri: the C2 recipe at N = 500 with leaf 32: positions after 100
  steps of the reference's "Rk4" (Dormand-Prince 5(4), dt 0.002, threshold
  1e-3), every stage rebuilding the tree (601 oracle matvecs; N = 500 keeps
  that within minutes of CPU time).
* c3: the C3 blob recipe at N = 20,000 (4 blobs, sigma 15), leaf 256, eps
  1e-4, max_iter 40, max_relative_coord 2048, smooth(0): lists, ranks,
  velocities.
* c4: the C4 recipe at 200 filaments x 50 beads (N = 10,000), bonded Morse
  (D 1, alpha 1, r0 2.2) plus gravity, leaf 256, eps 1e-4, max_iter 32,
  smooth(0): forces, lists, ranks, velocities.
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hignn_amd import weights, workloads  # noqa: E402
from oracle import forcing as OF  # noqa: E402
from oracle import integrators as OI  # noqa: E402
from oracle import oracle as O  # noqa: E402


def digest(x):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(x, np.float32).tobytes()).digest(),
                         dtype=np.uint8)


# name -> (positions, leaf, epsilon, max_iter, max_relative_coord, weights)
def setup(name):
    if name == "c1":
        return workloads.uniform_cube(2000, 0.05, 0), 64, 1e-5, 100, 500000, ("xavier", 0)
    if name == "c2":
        return workloads.uniform_cube(20000, 0.10, 1), 256, 1e-4, 32, 500000, ("smooth", 0)
    if name == "c2_dopri":
        return workloads.uniform_cube(500, 0.10, 1), 32, 1e-4, 32, 500000, ("smooth", 0)
    if name == "c3":
        return workloads.positions("c3_blobs_1m", n=20000), 256, 1e-4, 40, 2048, ("smooth", 0)
    if name == "c4":
        return workloads.filaments(200, 50, 2.2, 0.08, 3), 256, 1e-4, 32, 500000, ("smooth", 0)
    raise KeyError(name)


def weight_set(spec):
    kind, seed = spec
    return weights.xavier(seed) if kind == "xavier" else weights.smooth(seed)


def forces(name, x):
    if name == "c4":
        return OF.velocity_force(x, b=(0, 0, -1), bonds=OF.bond_forces(x, 200, 50, 1.0, 1.0, 2.2))
    return workloads.gravity(x.shape[0])


def oracle_velocity(x, leaf, eps, mi, mrc, w, f):
    m = O.OracleModel(x, leaf, epsilon=eps, max_iter=mi, max_relative_coord=mrc)
    m.update_coord()
    return m.dot(w, f)


def make(name):
    t0 = time.time()
    x, leaf, eps, mi, mrc, spec = setup(name)
    w = weight_set(spec)
    out = dict(x_sha256=digest(x), leaf=np.int64(leaf), epsilon=np.float64(eps),
               max_iter=np.int64(mi), max_relative_coord=np.int64(mrc),
               weights=np.array(f"{spec[0]}:{spec[1]}"))
    if name == "c2_dopri":
        def vel(t, xx):
            return oracle_velocity(xx, leaf, eps, mi, mrc, w, workloads.gravity(xx.shape[0]))
        xf, steps, sizes, calls = OI.dopri(x, vel, 0.002, 0.2)
        out.update(x_final=xf, calls=np.int64(calls), dt=np.float64(0.002), steps=np.int64(steps),
                   sizes=np.asarray(sizes, np.float64))
    else:
        f = forces(name, x)
        m = O.OracleModel(x, leaf, epsilon=eps, max_iter=mi, max_relative_coord=mrc)
        m.update_coord()
        e = m.export()
        out.update(f=f, u=m.dot(w, f), ranks=m.far_ranks(), perm=e["perm"], tree=e["tree"],
                   close=e["close"], far=e["far"], leaves=e["leaves"])
        if name == "c1":
            def vel(t, xx):
                return oracle_velocity(xx, leaf, eps, mi, mrc, w, workloads.gravity(xx.shape[0]))
            xf, calls = OI.euler(x, vel, 0.005, 0.05)
            out.update(x_euler=xf, euler_calls=np.int64(calls))
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    print(f"{name}: {os.path.getsize(path)} bytes in {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    for nm in sys.argv[1:] or ["c1", "c2", "c2_dopri", "c3", "c4"]:
        make(nm)
