"""Pins of oracle/forcing.py (CPU only): closed forms, invariants, an
independent neighbour search and an independent image layout."""
import numpy as np
import pytest

from oracle import forcing as F

f32 = np.float32


def test_tension_two_particles_closed_form():
    # stretched bond of length 3 with rest length 2: each end pulled towards the other by k_t (3 - 2)
    x = np.array([[0, 0, 0], [3, 0, 0]], f32)
    bend, tens = F.filament_terms(x, 1, 2, rest=2.0, k_t=5.0, k_b=1.0)
    np.testing.assert_array_equal(tens, [[5, 0, 0], [-5, 0, 0]])
    np.testing.assert_array_equal(bend, np.zeros((2, 3)))
    # compressed bond pushes apart
    x = np.array([[0, 0, 0], [0, 1, 0]], f32)
    _, tens = F.filament_terms(x, 1, 2, rest=2.0, k_t=1.0, k_b=1.0)
    np.testing.assert_array_equal(tens, [[0, -1, 0], [0, 1, 0]])


def test_bending_right_angle_closed_form():
    # arms (1,0,0) and (0,1,0): sin = 1, torque -(0,0,1); f21 = (0,-1,0), f23 = (1,0,0)
    x = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], f32)
    bend, _ = F.filament_terms(x, 1, 3, rest=1.0, k_t=0.0, k_b=2.0)
    np.testing.assert_array_equal(bend, [[0, -2, 0], [-2, 2, 0], [2, 0, 0]])


def test_straight_chain_has_no_bending():
    x = np.zeros((10, 3), f32)
    x[:, 2] = np.arange(10) * 1.5
    bend, tens = F.filament_terms(x, 1, 10, rest=1.5, k_t=3.0, k_b=7.0)
    assert not bend.any() and not tens.any()


def chains(rng, n_fil, n_chain, step=2.0):
    x = []
    for c in range(n_fil):
        d = rng.normal(size=(n_chain, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        x.append(rng.uniform(0, 50, 3) + np.cumsum(step * d, axis=0))
    return np.concatenate(x).astype(f32)


def test_filament_forces_newton_and_rotation():
    rng = np.random.default_rng(1)
    x = chains(rng, 3, 12)
    bend, tens = F.filament_terms(x, 3, 12, rest=1.8, k_t=4.0, k_b=2.5)
    for t in (bend, tens):
        for c in range(3):
            assert np.abs(t[12 * c:12 * c + 12].sum(axis=0)).max() < 1e-5 * np.abs(t).max()
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    xr = (x.astype(np.float64) @ q.T).astype(f32)
    bend_r, tens_r = F.filament_terms(xr, 3, 12, rest=1.8, k_t=4.0, k_b=2.5)
    np.testing.assert_allclose(bend_r, bend @ q.T, atol=2e-5 * np.abs(bend).max())
    np.testing.assert_allclose(tens_r, tens @ q.T, atol=2e-5 * np.abs(tens).max())


def test_morse_is_minus_gradient_of_potential():
    rng = np.random.default_rng(2)
    r = rng.normal(size=(50, 3))
    r *= (rng.uniform(0.5, 7.0, 50) / np.linalg.norm(r, axis=1))[:, None]
    r = r.astype(f32)
    got = F.morse_pair(r)
    rn = np.linalg.norm(r.astype(np.float64), axis=1)
    h = 1e-6
    dU = (F.morse_potential(rn + h) - F.morse_potential(rn - h)) / (2 * h)
    want = -dU[:, None] * r / rn[:, None]
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6)
    # zero force at the equilibrium distance, attraction beyond, repulsion inside
    assert np.abs(F.morse_pair(np.array([[2.5, 0, 0]], f32))).max() < 1e-6
    assert F.morse_pair(np.array([[4.0, 0, 0]], f32))[0, 0] < 0
    assert F.morse_pair(np.array([[1.5, 0, 0]], f32))[0, 0] > 0
    assert not F.morse_pair(np.array([[1e-4, 0, 0]], f32)).any()


def test_neighbour_search_against_kdtree():
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(3)
    x = rng.uniform(0, 20, size=(400, 3)).astype(f32)
    nb = F.neighbours(x, x, 3.0)
    tree = cKDTree(x.astype(np.float64))
    for i, got in enumerate(nb):
        want = sorted(tree.query_ball_point(x[i].astype(np.float64), 3.0))
        assert sorted(got.tolist()) == want


def test_morse_force_non_periodic_newton():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 15, size=(300, 3)).astype(f32)
    f = F.morse_force(x, 4.0)
    assert np.abs(f.sum(axis=0)).max() < 1e-4 * np.abs(f).max()
    # the same force from explicit pair sums
    i, j = 7, None
    d = np.linalg.norm(x.astype(np.float64) - x[i], axis=1)
    js = np.nonzero((d < 4.0) & (np.arange(300) != i))[0]
    want = F.morse_pair((x[i] - x[js]).astype(f32)).sum(axis=0)
    np.testing.assert_allclose(f[i], want, rtol=1e-12)


def images_by_bits(x, eps, lo, hi):
    """Image n of a target carries the shift of its k-th duplicated axis when
    bit (num >> (k+1)) of n is set."""
    src, tgt = [], []
    for i, p in enumerate(np.asarray(x, f32)):
        dup = [(-2 if p[j] < f32(lo[j] + f32(eps)) else (2 if p[j] > f32(hi[j] - f32(eps)) else 1))
               for j in range(3)]
        num = 2 ** sum(d != 1 for d in dup)
        for n in range(num):
            q, k = p.copy(), 0
            for j in range(3):
                if dup[j] != 1:
                    if n & (num >> (k + 1)):
                        L = f32(f32(hi[j]) - f32(lo[j]))
                        q[j] = f32(q[j] + (-L if dup[j] == 2 else L))
                    k += 1
            src.append(q)
            tgt.append(i)
    return np.array(src, f32), np.array(tgt)


def test_source_images_layout():
    rng = np.random.default_rng(5)
    lo, hi = np.array([0, -5, 2], f32), np.array([20, 15, 30], f32)
    x = rng.uniform(lo, hi, size=(200, 3)).astype(f32)
    x[0] = [0.5, -4.5, 2.5]  # a corner: 8 images
    src, tgt = F.source_images(x, 3.0, lo, hi)
    s2, t2 = images_by_bits(x, 3.0, lo, hi)
    np.testing.assert_array_equal(src, s2)
    np.testing.assert_array_equal(tgt, t2)
    assert (tgt == 0).sum() == 8
    # every image of a particle differs from it by whole domain lengths
    L = hi - lo
    k = (src - x[tgt]) / L
    np.testing.assert_allclose(k, np.round(k), atol=1e-5)


def test_periodic_morse_reads_source_array_at_target_index():
    # one pair across the x boundary: target 0 near lo, target 1 near hi
    lo, hi = np.array([0, 0, 0], f32), np.array([10, 10, 10], f32)
    x = np.array([[0.5, 5, 5], [9.5, 5, 5]], f32)
    src, tgt = F.source_images(x, 2.0, lo, hi)
    # target 0 has an image at +10, target 1 at -10; sources: [x0, x0+10, x1, x1-10]
    np.testing.assert_array_equal(tgt, [0, 0, 1, 1])
    f = F.morse_force(x, 2.0, periodic=True, lo=lo, hi=hi)
    # neighbour of target 0 is source 3 (target 1); the displacement is taken
    # against source entry 1 (x0 + 10), as GetPotentialForce indexes it
    np.testing.assert_array_equal(f[0], F.morse_pair((x[0] - src[1])[None].astype(f32))[0])
    np.testing.assert_array_equal(f[1], F.morse_pair((x[1] - src[0])[None].astype(f32))[0])


def test_bond_forces_are_minus_the_energy_gradient():
    """Bonded Morse forces against central differences of the total bond
    energy, Newton's third law per chain, and zero force at the rest length."""
    from hignn_amd import workloads
    x = workloads.filaments(3, 6, 2.2, 0.08, 5).astype(np.float64)
    x += np.random.default_rng(0).normal(scale=0.3, size=x.shape)
    Fb = F.bond_forces(x.astype(np.float32), 3, 6, D=1.3, alpha=0.9, r0=2.1)
    xs = x.astype(np.float32).astype(np.float64)
    h = 1e-6
    for i in range(x.shape[0]):
        for j in range(3):
            xp, xm = xs.copy(), xs.copy()
            xp[i, j] += h
            xm[i, j] -= h
            g = (F.bond_energy(xp, 3, 6, 1.3, 0.9, 2.1) - F.bond_energy(xm, 3, 6, 1.3, 0.9, 2.1)) / (2 * h)
            assert abs(Fb[i, j] + g) < 1e-6, (i, j, Fb[i, j], -g)
    np.testing.assert_allclose(Fb.reshape(3, 6, 3).sum(axis=1), 0.0, atol=1e-12)
    # a straight chain at the rest length feels no bond force
    st = workloads.filaments(2, 5, 2.2, 0.08, 1)
    assert np.abs(F.bond_forces(st, 2, 5)).max() < 1e-6
