"""numpy restatement of the forcing terms around the matvec (TEST INFRASTRUCTURE).

Only tests/ import this module, as the checker of the device forcing kernels.

- ``filament_terms``: chain bending (python/simulate.py:14-53) and chain
  tension (python/simulate.py:55-88): float32 arithmetic per bond / triple,
  accumulated into float64 as numpy's ``F += float32 term`` does.
- ``velocity_force``: the force array of velocity_update
  (python/simulate.py:90-144): ``force = 0 (float32); force += b;
  force += bending; force += tension``.
- ``morse_pair``: CalculatePotentialForce (include/PotentialForce.hpp:115-140),
  De = 1, a = 1, re = 2.5.
- ``bond_forces``: the same Morse form with the bond parameters (D, alpha,
  r0) between consecutive beads of each chain (the filament workload's
  bonds), float64.
- ``source_images``: BuildSourceSites (include/PotentialForce.hpp:247-392),
  written as its loop over duplicated axes with halving block strides.
- ``morse_force``: GetPotentialForce (include/PotentialForce.hpp:468-580) with
  an O(n m) radius search whose test is the kd-tree's
  (include/PointCloudSearch.hpp:386-400, include/nanoflann.hpp:387-393,
  1546-1552): d = sqrt(sum (s - q)^2) in double, accepted when
  d * d < (double)(float)(eps * eps).  The neighbours are mapped to target
  indices (PotentialForce.hpp:424-432), the first entry equal to the target is
  dropped (:434-446), and each remaining neighbour t contributes
  Morse(x_i - S[t]) with S the source array (:548) when |r| > 1e-3.

Pinned by tests/test_forcing_oracle.py: closed forms (two-particle tension,
right-angle bending, zero bending of a straight chain), Newton's third law,
rotation covariance, the Morse force and the bond forces as minus the
finite-difference gradient of their potentials, the brute-force neighbour set against scipy's
cKDTree, and the image layout against an independent bit-by-bit reading.
"""
import numpy as np

f32 = np.float32


def _bond_nodal(a, b, k_t, rest):
    """chain_tension per bond: r = a - b, nodal = (r / |r|) * k_t (|r| - rest), float32."""
    r = (a - b).astype(f32)
    rn = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]).astype(f32) + r[:, 2] * r[:, 2]).astype(f32)
    fm = (f32(k_t) * (rn - f32(rest))).astype(f32)
    return ((r / rn[:, None]).astype(f32) * fm[:, None]).astype(f32)


def _cross(a, b):
    return np.stack([(a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1]),
                     (a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2]),
                     (a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0])], axis=1).astype(f32)


def _unit(a):
    n = np.sqrt((a[:, 0] * a[:, 0] + a[:, 1] * a[:, 1]).astype(f32) + a[:, 2] * a[:, 2]).astype(f32)
    return (a / n[:, None]).astype(f32)


def _triple_forces(x0, x1, x2):
    """chain_bending per triple: (f21, f23) before the k_b factor, float32."""
    a1 = _unit((x1 - x0).astype(f32))
    a2 = _unit((x2 - x1).astype(f32))
    c = ((a1[:, 0] * a2[:, 0] + a1[:, 1] * a2[:, 1]).astype(f32) + a1[:, 2] * a2[:, 2]).astype(f32)
    s2 = (f32(1) - c * c).astype(f32)
    s2 = np.where(s2 < 0, f32(0), s2)
    s = np.sqrt(s2).astype(f32)
    tq = ((-_cross(a1, a2)) * s[:, None]).astype(f32)
    return _cross(tq, a1), _cross(tq, a2)


def filament_terms(x, n_fil, n_chain, rest, k_t, k_b):
    """(bending, tension) float64 arrays (n,3); zero outside the chains."""
    x = np.asarray(x, dtype=f32)
    n = x.shape[0]
    bend = np.zeros((n, 3))
    tens = np.zeros((n, 3))
    for c in range(n_fil):
        lo = c * n_chain
        p = x[lo:lo + n_chain]
        if n_chain >= 3:
            f21, f23 = _triple_forces(p[:-2], p[1:-1], p[2:])
            B = np.zeros((n_chain, 3))
            B[:-2] += f21
            B[1:-1] -= (f21 + f23).astype(f32)
            B[2:] += f23
            bend[lo:lo + n_chain] = float(k_b) * B
        if n_chain >= 2:
            nodal = _bond_nodal(p[:-1], p[1:], k_t, rest)
            T = np.zeros((n_chain, 3))
            T[:-1] -= nodal
            T[1:] += nodal
            tens[lo:lo + n_chain] = T
    return bend, tens


def bond_forces(x, n_fil, n_chain, D=1.0, alpha=1.0, r0=2.2):
    """Morse bonds between consecutive beads of n_fil chains of n_chain:
    U(r) = D (1 - exp(-alpha (r - r0)))^2 per bond (the Morse form of
    CalculatePotentialForce, include/PotentialForce.hpp:115-140, with the
    bond's parameters), float64 (n,3) from float32 positions; particles past
    the chains get zero."""
    x = np.asarray(x, dtype=f32).astype(np.float64)
    out = np.zeros_like(x)
    m = n_fil * n_chain
    X = x[:m].reshape(n_fil, n_chain, 3)
    d = X[:, 1:] - X[:, :-1]
    r = np.sqrt(np.sum(d * d, axis=2, keepdims=True))
    e = np.exp(-alpha * (r - r0))
    fb = 2.0 * D * alpha * e * (1.0 - e) / r * d  # on bead k of bond (k, k+1)
    F = np.zeros_like(X)
    F[:, :-1] += fb
    F[:, 1:] -= fb
    out[:m] = F.reshape(m, 3)
    return out


def bond_energy(x, n_fil, n_chain, D=1.0, alpha=1.0, r0=2.2):
    """Total bond energy sum_bonds D (1 - exp(-alpha (r - r0)))^2 (float64)."""
    X = np.asarray(x, dtype=np.float64)[:n_fil * n_chain].reshape(n_fil, n_chain, 3)
    r = np.linalg.norm(X[:, 1:] - X[:, :-1], axis=2)
    return float(np.sum(D * (1.0 - np.exp(-alpha * (r - r0))) ** 2))


def velocity_force(x, b=(0, 0, 0), n_fil=0, n_chain=0, rest=0.0, k_t=0.0, k_b=0.0, morse=None,
                   bonds=None):
    """The float32 force of velocity_update; ``bonds`` and then ``morse``
    (float64 (n,3)) are added last when given, as the device forcing does."""
    x = np.asarray(x, dtype=f32)
    f = np.zeros(x.shape, dtype=f32)
    f += np.asarray(b, dtype=f32)
    if n_fil > 0 and n_chain > 1:
        bend, tens = filament_terms(x, n_fil, n_chain, rest, k_t, k_b)
        f = (f.astype(np.float64) + bend).astype(f32)
        f = (f.astype(np.float64) + tens).astype(f32)
    if bonds is not None:
        f = (f.astype(np.float64) + bonds).astype(f32)
    if morse is not None:
        f = (f.astype(np.float64) + morse).astype(f32)
    return f


def morse_pair(r):
    """Force for displacements r (m,3) float32 -> float64 (m,3); zero where |r| <= 1e-3."""
    r = np.asarray(r, dtype=f32)
    rn = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]).astype(f32) + r[:, 2] * r[:, 2]).astype(f32)
    d = (rn - f32(2.5)).astype(f32)
    fmag = -2.0 * (np.exp(-d.astype(np.float64)) - np.exp(-2.0 * d.astype(np.float64)))
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = (fmag / rn.astype(np.float64)).astype(f32)
        out = (r * ratio[:, None]).astype(f32).astype(np.float64)
    out[~(rn.astype(np.float64) > 1e-3)] = 0.0
    return out


def morse_potential(rn):
    """U(r) = De (1 - exp(-a (r - re)))^2, whose -dU/dr r-hat is morse_pair."""
    return (1.0 - np.exp(-(np.asarray(rn, np.float64) - 2.5))) ** 2


def source_images(x, eps, lo, hi):
    """BuildSourceSites with a periodic domain: (sources (m,3) float32,
    target index of each source)."""
    x = np.asarray(x, dtype=f32)
    eps = f32(eps)
    lo = np.asarray(lo, dtype=f32)
    hi = np.asarray(hi, dtype=f32)
    core_lo = (lo + eps).astype(f32)
    core_hi = (hi - eps).astype(f32)
    size = (hi - lo).astype(f32)
    src, tgt = [], []
    for i in range(x.shape[0]):
        kind = []
        for j in range(3):
            if x[i, j] < core_lo[j]:
                kind.append(-2)
            elif x[i, j] > core_hi[j]:
                kind.append(2)
            else:
                kind.append(1)
        num = int(np.prod([2 if k != 1 else 1 for k in kind]))
        off = np.zeros((num, 3), dtype=f32)
        stride = num
        for j in range(3):
            if kind[j] == 1:
                continue
            shift = -size[j] if kind[j] == 2 else size[j]
            for m0 in range(0, num, stride):
                off[m0 + stride // 2:m0 + stride, j] = shift
            stride //= 2
        for q in range(num):
            src.append((x[i] + off[q]).astype(f32))
            tgt.append(i)
    return np.array(src, dtype=f32).reshape(-1, 3), np.array(tgt, dtype=np.int64)


def neighbours(q, src, eps):
    """Indices of sources within the kd-tree's radius test of each query."""
    e2 = float(f32(eps) * f32(eps))
    s = np.asarray(src, dtype=np.float64)
    out = []
    for i in range(q.shape[0]):
        d2 = ((s - np.asarray(q[i], dtype=np.float64)) ** 2).sum(axis=1)
        d = np.sqrt(d2)
        out.append(np.nonzero(d * d < e2)[0])
    return out


def morse_force(x, eps, periodic=False, lo=None, hi=None):
    """GetPotentialForce: float64 (n,3) sums of float32 pair forces."""
    x = np.asarray(x, dtype=f32)
    if periodic:
        src, tgt = source_images(x, eps, lo, hi)
    else:
        src, tgt = x, np.arange(x.shape[0])
    nb = neighbours(x, src, eps)
    f = np.zeros((x.shape[0], 3))
    for i in range(x.shape[0]):
        t = tgt[nb[i]]
        hit = np.nonzero(t == i)[0]
        if hit.size:
            t = np.delete(t, hit[0])
        if t.size == 0:
            continue
        r = (x[i][None, :] - src[t]).astype(f32)
        f[i] = morse_pair(r).sum(axis=0)
    return f
