"""Synthetic particle configurations and forcing for the benchmark suite.

Sphere radius a = 1.  A cube of side L holding N spheres at volume fraction
phi has L = (N * 4/3 pi / phi)^(1/3).

Placement is rejection sampling in rounds: each round draws a batch of
candidates, drops every candidate closer than ``dmin`` to an accepted sphere,
and of two candidates of the same round that clash keeps the earlier one.
Candidates come from numpy's PCG64 seeded per configuration, so a
configuration is a pure function of its arguments.

Forcing follows python/simulate.py:90-144 (uniform body force, plus chain
tension python/simulate.py:57-88 for filaments).  The Morse bond force of the
filament workload is written out here (D, alpha, r0).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.spatial import cKDTree


def cube_side(n: int, phi: float) -> float:
    return (n * 4.0 / 3.0 * math.pi / phi) ** (1.0 / 3.0)


def _reject_sample(n, draw, dmin, rng, max_rounds=10000):
    acc = np.zeros((0, 3), dtype=np.float64)
    rounds = 0
    while acc.shape[0] < n:
        rounds += 1
        if rounds > max_rounds:
            raise RuntimeError("rejection sampling did not converge")
        need = n - acc.shape[0]
        cand = draw(rng, max(64, int(need * 1.25)))
        if acc.shape[0]:
            d, _ = cKDTree(acc).query(cand, k=1, distance_upper_bound=dmin)
            cand = cand[~(d < dmin)]
        if cand.shape[0] == 0:
            continue
        pairs = cKDTree(cand).query_pairs(dmin, output_type="ndarray")
        if pairs.size:
            bad = np.zeros(cand.shape[0], dtype=bool)
            bad[np.maximum(pairs[:, 0], pairs[:, 1])] = True
            cand = cand[~bad]
        acc = np.concatenate([acc, cand[:need]], axis=0)
    return acc


def uniform_cube(n: int, phi: float, seed: int, dmin: float = 2.05) -> np.ndarray:
    """N spheres uniform in a cube [0, L)^3, pairwise centre distance >= dmin."""
    L = cube_side(n, phi)
    rng = np.random.Generator(np.random.PCG64(seed))
    pts = _reject_sample(n, lambda g, m: g.uniform(0.0, L, size=(m, 3)), dmin, rng)
    return np.ascontiguousarray(pts.astype(np.float32))


def gaussian_blobs(n: int, nblobs: int, sigma: float, phi: float, seed: int,
                   dmin: float = 2.05) -> np.ndarray:
    """N spheres in ``nblobs`` Gaussian blobs (std ``sigma``) whose centres are
    uniform in the cube sized for mean volume fraction ``phi``."""
    L = cube_side(n, phi)
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(0.0, L, size=(nblobs, 3))

    def draw(g, m):
        k = g.integers(0, nblobs, size=m)
        return centres[k] + g.normal(0.0, sigma, size=(m, 3))

    pts = _reject_sample(n, draw, dmin, rng)
    return np.ascontiguousarray(pts.astype(np.float32))


def filaments(nfil: int, nbead: int, bond: float, phi: float, seed: int) -> np.ndarray:
    """Straight chains of ``nbead`` beads (spacing ``bond``) in random
    directions, chain centres uniform in the cube sized for the total bead
    count at volume fraction ``phi``.  Bead k of filament f is row f*nbead+k."""
    n = nfil * nbead
    L = cube_side(n, phi)
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(0.0, L, size=(nfil, 1, 3))
    d = rng.normal(size=(nfil, 1, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    s = (np.arange(nbead) - 0.5 * (nbead - 1)) * bond
    pts = centres + d * s[None, :, None]
    return np.ascontiguousarray(pts.reshape(n, 3).astype(np.float32))


def filament_cloud(nfil: int = 1024, nchain: int = 21, rest: float = 3.0, radius: float = 240.0,
                   seed: int = 0) -> np.ndarray:
    """Straight filaments of ``nchain`` beads at spacing ``rest`` in random
    directions, centres uniform in the ball of the given radius (less half a
    filament length) -- the shape of the reference's filament_cloud
    configuration (python/config/filament_cloud.json: 1024 x 21, rest length
    3, R2 = 240).  Bead k of filament f is row f*nchain+k."""
    rng = np.random.Generator(np.random.PCG64(seed))
    half = 0.5 * (nchain - 1) * rest
    r = (radius - half) * rng.uniform(0.0, 1.0, size=nfil) ** (1.0 / 3.0)
    c = rng.normal(size=(nfil, 3))
    c *= (r / np.linalg.norm(c, axis=1))[:, None]
    d = rng.normal(size=(nfil, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = (np.arange(nchain) - 0.5 * (nchain - 1)) * rest
    pts = c[:, None, :] + d[:, None, :] * s[None, :, None]
    return np.ascontiguousarray(pts.reshape(-1, 3).astype(np.float32))


def gravity(n: int, g=(0.0, 0.0, -1.0)) -> np.ndarray:
    f = np.empty((n, 3), dtype=np.float32)
    f[:] = np.asarray(g, dtype=np.float32)
    return f


def morse_bonds(x: np.ndarray, nfil: int, nbead: int, D=1.0, alpha=1.0, r0=2.2) -> np.ndarray:
    """Forces of Morse bonds U(r) = D (1 - exp(-alpha (r - r0)))^2 between
    consecutive beads of each filament."""
    X = np.asarray(x, dtype=np.float64).reshape(nfil, nbead, 3)
    d = X[:, 1:, :] - X[:, :-1, :]
    r = np.linalg.norm(d, axis=2, keepdims=True)
    e = np.exp(-alpha * (r - r0))
    dU = 2.0 * D * alpha * e * (1.0 - e)  # dU/dr
    fb = dU * d / r  # force on bead k from bond (k, k+1), pulling towards k+1
    F = np.zeros_like(X)
    F[:, :-1, :] += fb
    F[:, 1:, :] -= fb
    return F.reshape(-1, 3).astype(np.float32)


# Named configurations (sizes, seeds, solver parameters).
CONFIGS = {
    "c1_uniform_2k": dict(n=2000, phi=0.05, seed=0, leaf=64, epsilon=1e-5, max_iter=50),
    "c2_uniform_100k": dict(n=100_000, phi=0.10, seed=1, leaf=256, epsilon=1e-4, max_iter=32),
    "c3_blobs_1m": dict(n=1_000_000, phi=0.10, seed=2, nblobs=200, sigma=15.0, leaf=256,
                        epsilon=1e-4, max_iter=32),
    "c4_filaments_100k": dict(nfil=2000, nbead=50, bond=2.2, phi=0.08, seed=3, leaf=256,
                              epsilon=1e-4, max_iter=32),
    "c5_instances_64": dict(count=64, nmin=5000, nmax=200_000, seed=4, phi=(0.05, 0.15)),
}


def instances(count: int = 64, nmin: int = 5000, nmax: int = 200_000, seed: int = 4,
              phi_range=(0.05, 0.15)):
    """Independent suspensions of the many-instance workload: N log-uniform
    in [nmin, nmax] and volume fraction uniform in phi_range per instance
    (drawn from one PCG64 stream seeded with ``seed``), each a uniform cube
    with its own seed (seed * 1000 + k).  Returns (positions of all instances
    concatenated, per-instance counts)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), size=count)).astype(np.int64)
    phis = rng.uniform(phi_range[0], phi_range[1], size=count)
    xs = [uniform_cube(int(ns[k]), float(phis[k]), seed * 1000 + k) for k in range(count)]
    return np.ascontiguousarray(np.concatenate(xs, axis=0)), ns


def positions(name: str, n: int | None = None) -> np.ndarray:
    """Positions of a named configuration; ``n`` overrides N for scaled-down
    runs of the same recipe (same seed, same volume fraction)."""
    c = CONFIGS[name]
    if name.startswith("c3"):
        nn = n or c["n"]
        nb = max(1, int(round(c["nblobs"] * nn / c["n"])))
        return gaussian_blobs(nn, nb, c["sigma"], c["phi"], c["seed"])
    if name.startswith("c4"):
        nfil = c["nfil"] if n is None else max(1, n // c["nbead"])
        return filaments(nfil, c["nbead"], c["bond"], c["phi"], c["seed"])
    return uniform_cube(n or c["n"], c["phi"], c["seed"])
