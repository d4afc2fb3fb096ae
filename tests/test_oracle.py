"""Pins of the CPU oracle (oracle/hh_oracle.cpp) that do not use the oracle's
own formulas: closed forms, an independent numpy evaluation, brute force on
small inputs, invariants of the tree and the block partition, and the
cross-interpolation property of the ACA."""
import os

import numpy as np
import pytest

from hignn_amd import weights, workloads
from oracle import oracle as O

REF_PKL = "/root/reference/nn/two_body_unbounded.pkl"


def np_mlp(w, r):
    """numpy float64 evaluation of Net.forward (python/convert.py:11-20)."""
    p = weights.unpack(w)
    r = np.asarray(r, dtype=np.float64)
    h1 = np.tanh(r @ p["W1"].T.astype(np.float64) + p["b1"])
    h2 = np.tanh(h1 @ p["W2"].T.astype(np.float64) + p["b2"])
    y = h2 @ p["W3"].T.astype(np.float64) + p["b3"]
    y[:, [0, 4, 8]] += 1.0
    nr = np.linalg.norm(r, axis=1, keepdims=True)
    with np.errstate(divide="ignore", invalid="ignore"):
        y = y / nr
    y[nr[:, 0] < 1e-5] = np.eye(3).reshape(-1)
    return y


def rand_disp(n, seed, lo=1e-3, hi=1e3):
    g = np.random.default_rng(seed)
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d * np.exp(g.uniform(np.log(lo), np.log(hi), size=n))[:, None]


# ---------------------------------------------------------------- pair function
def test_mlp_identity_below_cutoff(xavier_w):
    r = np.array([[0, 0, 0], [9e-6, 0, 0], [0, 5e-6, 5e-6]], dtype=np.float64)
    for out in (O.mlp_f64(xavier_w, r), O.mlp_f32(xavier_w, r)):
        np.testing.assert_array_equal(out, np.tile(np.eye(3).reshape(-1), (3, 1)))


def test_mlp_closed_form_w2_zero(xavier_w):
    p = weights.unpack(xavier_w)
    p = {k: v.copy() for k, v in p.items()}
    p["W2"][:] = 0.0
    w = weights.pack(p)
    r = rand_disp(200, 1)
    y = p["W3"].astype(np.float64) @ np.tanh(p["b2"].astype(np.float64)) + p["b3"]
    y = y + np.eye(3).reshape(-1)
    expect = y[None, :] / np.linalg.norm(r, axis=1)[:, None]
    np.testing.assert_allclose(O.mlp_f64(w, r), expect, rtol=1e-12, atol=0)


def test_mlp_closed_form_oseen_scaling(xavier_w):
    """W3 = b3 = 0 leaves M = I/|r| exactly: checks the +I, /|r| and the
    row-major 3x3 layout."""
    p = {k: v.copy() for k, v in weights.unpack(xavier_w).items()}
    p["W3"][:] = 0.0
    p["b3"][:] = 0.0
    w = weights.pack(p)
    r = rand_disp(100, 2)
    expect = np.eye(3).reshape(-1)[None, :] / np.linalg.norm(r, axis=1)[:, None]
    np.testing.assert_allclose(O.mlp_f64(w, r), expect, rtol=1e-15)


def test_mlp_matches_numpy(xavier_w):
    r = rand_disp(500, 3)
    ref = np_mlp(xavier_w, r)
    np.testing.assert_allclose(O.mlp_f64(xavier_w, r), ref, rtol=1e-11, atol=1e-14)
    got32 = O.mlp_f32(xavier_w, r.astype(np.float32)).astype(np.float64)
    ref32 = np_mlp(xavier_w, r.astype(np.float32).astype(np.float64))
    err = np.abs(got32 - ref32).max(axis=1) / np.abs(ref32).max(axis=1)
    assert err.max() < 5e-6


@pytest.mark.skipif(not os.path.exists(REF_PKL), reason="reference weights not present")
def test_trained_weights_properties():
    """Measured properties of the shipped trained net: blocks symmetric to
    ~1e-8, and the Oseen tensor 3/(4r)(I + rr^T/r^2) within a few percent
    at r >= 10."""
    w = weights.load_pkl(REF_PKL)
    r = rand_disp(400, 4, lo=10.0, hi=500.0)
    M = O.mlp_f64(w, r).reshape(-1, 3, 3)
    asym = np.linalg.norm(M - M.transpose(0, 2, 1)) / np.linalg.norm(M)
    assert asym < 1e-6
    nr = np.linalg.norm(r, axis=1)
    rh = r / nr[:, None]
    oseen = 0.75 / nr[:, None, None] * (np.eye(3)[None] + rh[:, :, None] * rh[:, None, :])
    dev = np.linalg.norm(M - oseen, axis=(1, 2)) / np.linalg.norm(oseen, axis=(1, 2))
    assert np.median(dev) < 0.02


# ---------------------------------------------------------------- tree
def test_principal_axis_matches_eigh():
    g = np.random.default_rng(5)
    for _ in range(50):
        a = g.normal(size=(3, 3))
        c = a @ a.T
        cov6 = np.array([c[0, 0], c[1, 1], c[2, 2], c[0, 1], c[1, 2], c[2, 0]])
        u = O.principal_axis(cov6)
        ev, V = np.linalg.eigh(c)
        v = V[:, -1]
        assert abs(abs(u @ v) - 1.0) < 1e-10
        big = int(np.argmax(np.abs(u)))
        assert u[big] > 0


def tree_nodes(tree):
    """Reachable nodes in breadth-first order."""
    order, k = [0], 0
    while k < len(order):
        nd = order[k]
        k += 1
        if tree[nd, 0] != 0:
            order += [int(tree[nd, 0]), int(tree[nd, 1])]
    return order


@pytest.mark.parametrize("case", ["uniform", "blobs", "filaments", "lattice"])
def test_tree_invariants(case):
    if case == "uniform":
        x, leaf = workloads.uniform_cube(3000, 0.1, 11), 64
    elif case == "blobs":
        x, leaf = workloads.gaussian_blobs(4000, 3, 10.0, 0.1, 12), 128
    elif case == "filaments":
        x, leaf = workloads.filaments(60, 50, 2.2, 0.08, 13), 100
    else:
        x = (np.stack(np.meshgrid(*[np.arange(12.0) * 2.5] * 3), -1).reshape(-1, 3)).astype(np.float32)
        leaf = 50
    m = O.OracleModel(x, leaf)
    m.update_coord()
    e = m.export()
    tree, perm = e["tree"].astype(np.int64), e["perm"].astype(np.int64)
    n = x.shape[0]
    assert np.array_equal(np.sort(perm), np.arange(n))
    order = tree_nodes(tree)
    # rows 1 and 2 are the unreachable first division (one rank)
    assert 1 not in order and 2 not in order
    assert tree[0, 0] == 3 and tree[0, 1] == 4
    # breadth-first numbering: children appear in increasing id order
    kids = [int(tree[nd, c]) for nd in order if tree[nd, 0] for c in (0, 1)]
    assert kids == sorted(kids) and kids[0] == 3
    xp = x[perm].astype(np.float64)
    for nd in order:
        s, t = tree[nd, 2], tree[nd, 3]
        if tree[nd, 0] == 0:
            assert t - s <= leaf
            continue
        a, b = tree[nd, 0], tree[nd, 1]
        assert tree[a, 2] == s and tree[a, 3] == tree[b, 2] and tree[b, 3] == t
        assert tree[a, 4] == tree[nd, 4] + 1 == tree[b, 4]
        # split agrees with an independent PCA of the node's particles
        pts = xp[s:t]
        mu = pts.mean(axis=0)
        ev, V = np.linalg.eigh(np.cov((pts - mu).T, bias=True))
        if ev[-1] < ev[-2] * (1 + 1e-6):
            continue  # principal axis not unique (lattice symmetry)
        u = V[:, -1]
        if u[int(np.argmax(np.abs(u)))] < 0:
            u = -u
        proj = (pts - mu) @ u
        tol = 1e-9 * np.abs(proj).max()
        split = tree[a, 3] - s
        assert np.all(proj[:split] <= tol) and np.all(proj[split:] > -tol)


def test_tree_tiny_is_single_leaf():
    x = workloads.uniform_cube(40, 0.05, 3)
    m = O.OracleModel(x, 64)
    m.update_coord()
    e = m.export()
    assert e["tree"].shape == (3, 5)
    assert list(e["tree"][0]) == [0, 0, 0, 40, 0]
    assert e["close"].tolist() == [[0, 0]] and e["far"].shape[0] == 0


# ---------------------------------------------------------------- boxes, admissibility, lists
def np_is_far(a, b):
    a = a.astype(np.float32)
    b = b.astype(np.float32)
    if not (a[0] > b[1] or a[1] < b[0] or a[2] > b[3] or a[3] < b[2] or a[4] > b[5] or a[5] < b[4]):
        return False
    d0 = np.float32(0)
    d1 = np.float32(0)
    dist = np.float32(0)
    for j in range(3):
        t = a[2 * j] - a[2 * j + 1]
        d0 = np.float32(d0 + t * t)
        t = b[2 * j] - b[2 * j + 1]
        d1 = np.float32(d1 + t * t)
        t = a[2 * j] + a[2 * j + 1] - b[2 * j] - b[2 * j + 1]
        dist = np.float32(dist + t * t)
    dist = np.float32(dist * np.float32(0.25))
    return bool(dist > d0 and dist > d1)


@pytest.fixture(scope="module")
def lists_case():
    x = workloads.uniform_cube(2500, 0.08, 21)
    m = O.OracleModel(x, 64, use_symmetry=False)
    m.update_coord()
    ms = O.OracleModel(x, 64, use_symmetry=True)
    ms.update_coord()
    return x, m, m.export(), ms.export()


def test_boxes_brute_force(lists_case):
    x, m, e, _ = lists_case
    tree, perm = e["tree"].astype(np.int64), e["perm"].astype(np.int64)
    xp = x[perm]
    for nd in tree_nodes(tree):
        pts = xp[tree[nd, 2]:tree[nd, 3]]
        want = np.array([pts[:, 0].min(), pts[:, 0].max(), pts[:, 1].min(), pts[:, 1].max(),
                         pts[:, 2].min(), pts[:, 2].max()], dtype=np.float32)
        np.testing.assert_array_equal(e["aux"][nd], want)


def test_admissibility_brute_force(lists_case):
    x, m, e, _ = lists_case
    aux = e["aux"]
    for I, J in e["far"]:
        assert np_is_far(aux[I], aux[J])
    g = np.random.default_rng(0)
    nodes = tree_nodes(e["tree"].astype(np.int64))
    for _ in range(400):
        a, b = g.choice(nodes, 2)
        assert m.is_far(a, b) == np_is_far(aux[a], aux[b])


def coverage(e, n, sym):
    tree = e["tree"].astype(np.int64)
    cov = np.zeros((n, n), dtype=np.int32)
    for (I, J) in list(e["close"]) + list(e["far"]):
        a, b = slice(tree[I, 2], tree[I, 3]), slice(tree[J, 2], tree[J, 3])
        cov[a, b] += 1
        if sym and I != J:
            cov[b, a] += 1
    return cov


def test_blocks_tile_all_pairs_once(lists_case):
    x, _, e, es = lists_case
    n = x.shape[0]
    assert np.all(coverage(e, n, False) == 1)
    # with symmetry the J >= I half plus its transpose covers every pair once
    assert np.all(coverage(es, n, True) == 1)


# ---------------------------------------------------------------- dot parts
def test_close_dot_brute_force(xavier_w):
    x = workloads.uniform_cube(400, 0.08, 31)
    f = np.random.default_rng(1).normal(size=x.shape).astype(np.float32)
    m = O.OracleModel(x, 32)
    m.update_coord()
    e = m.export()
    tree, perm = e["tree"].astype(np.int64), e["perm"].astype(np.int64)
    xp, fp = x[perm].astype(np.float64), f[perm].astype(np.float64)
    u = np.zeros_like(xp)
    for I, J in e["close"]:
        ii = np.arange(tree[I, 2], tree[I, 3])
        jj = np.arange(tree[J, 2], tree[J, 3])
        r = (xp[jj][None, :, :] - xp[ii][:, None, :]).reshape(-1, 3)
        M = np_mlp(xavier_w, r.astype(np.float32).astype(np.float64)).reshape(len(ii), len(jj), 3, 3)
        u[ii] += np.einsum("ijrc,jc->ir", M, fp[jj])
        if J > I:  # symmetric reuse: same block, not transposed
            u[jj] += np.einsum("ijrc,ic->jr", M, fp[ii])
    want = np.zeros_like(u)
    want[perm] = u
    got = m.dot(xavier_w, f, parts=1)
    assert O.rel_l2(got, want) < 1e-6


def test_dense_dot_brute_force(xavier_w):
    x = workloads.uniform_cube(300, 0.08, 32)
    f = np.random.default_rng(2).normal(size=x.shape).astype(np.float32)
    m = O.OracleModel(x, 32)
    m.update_coord()
    r = (x[None, :, :].astype(np.float64) - x[:, None, :]).reshape(-1, 3)
    M = np_mlp(xavier_w, r.astype(np.float32).astype(np.float64)).reshape(300, 300, 3, 3)
    S = 0.5 * (M + M.transpose(0, 1, 3, 2))
    want = np.einsum("ijrc,jc->ir", S, f.astype(np.float64))
    assert O.rel_l2(m.dense_dot(xavier_w, f), want) < 1e-6


@pytest.fixture(scope="module")
def aca_case(xavier_w):
    x = workloads.uniform_cube(1500, 0.05, 41)
    m = O.OracleModel(x, 64, epsilon=1e-3, max_iter=40)
    m.update_coord()
    return x, m


def test_aca_cross_interpolation(aca_case, xavier_w):
    """The rank-k approximant reproduces the symmetrised block exactly on the
    pivot rows and the first k pivot columns."""
    x, m = aca_case
    e = m.export()
    tree, perm = e["tree"].astype(np.int64), e["perm"].astype(np.int64)
    xp = x[perm]
    sizes = [(tree[I, 3] - tree[I, 2]) * (tree[J, 3] - tree[J, 2]) for I, J in e["far"]]
    for b in np.argsort(sizes)[-3:]:
        fa = m.aca_factors(xavier_w, int(b), 40)
        I, J, k = fa["I"], fa["J"], fa["k"]
        xi, xj = xp[tree[I, 2]:tree[I, 3]], xp[tree[J, 2]:tree[J, 3]]
        r = (xj[None, :, :] - xi[:, None, :]).reshape(-1, 3)
        M = O.mlp_f32(xavier_w, r).astype(np.float64).reshape(len(xi), len(xj), 3, 3)
        S = M.copy()
        off = ~np.eye(3, dtype=bool)
        S[..., off] = 0.5 * (M + M.transpose(0, 1, 3, 2))[..., off]
        A = np.einsum("lirc,ljcd->ijrd", fa["C"], fa["Q"])
        scale = np.abs(S).max()
        rows = fa["ipiv"]
        cols = fa["jpiv"][:k]
        assert np.abs(A[rows] - S[rows]).max() < 1e-6 * scale
        assert np.abs(A[:, cols] - S[:, cols]).max() < 1e-6 * scale
        # the approximation error shrinks well below the block norm
        assert np.linalg.norm(A - S) < 0.1 * np.linalg.norm(S)


def test_dot_vs_fp64_blocks(aca_case, xavier_w):
    """Full oracle dot against its FP64 limit (exact far blocks, same block
    convention): the difference is the ACA truncation (epsilon 1e-3)."""
    x, m = aca_case
    f = workloads.gravity(x.shape[0])
    u = m.dot(xavier_w, f)
    tg = np.arange(0, x.shape[0], 7)
    ref = m.dense64_blocks(xavier_w, f, tg)
    err = O.rel_l2(u[tg], ref)
    assert err < 5e-3
    # and the near part alone is exact to float rounding
    un = m.dot(xavier_w, f, parts=1)
    assert np.isfinite(un).all()


# ------------------------------------------------ ACA decisions of the reference
def _sym_block(x, tree, perm, I, J, w):
    """The ACA's entries of far block (I, J): S(Net(x_j - x_i)) with the
    off-diagonal entries averaged (src/FarDot.cpp:441-456), from the FP32
    pair function, as float64 (m, n, 3, 3)."""
    xp = x[perm]
    xi, xj = xp[tree[I, 2]:tree[I, 3]], xp[tree[J, 2]:tree[J, 3]]
    r = (xj[None, :, :] - xi[:, None, :]).reshape(-1, 3)
    M = O.mlp_f32(w, r).reshape(len(xi), len(xj), 3, 3)
    # float sum, then the double 0.5 (0.5 * (float + float) in the reference)
    T = (M + M.transpose(0, 1, 3, 2)).astype(np.float64)
    S = M.astype(np.float64)
    off = ~np.eye(3, dtype=bool)
    S[..., off] = 0.5 * T[..., off]
    return S


def _pick(vals, eligible):
    """First index of the largest value among eligible indices (np.argmax
    keeps the first maximum), and whether the runner-up is so close that
    rounding of the test's own sums could reorder them."""
    v = np.where(eligible, vals, -np.inf)
    i = int(np.argmax(v))
    w = v.copy()
    w[i] = -np.inf
    close = np.isfinite(w).any() and (v[i] - w.max()) <= 1e-9 * np.abs(vals[eligible]).max()
    return i, close


@pytest.fixture(scope="module")
def aca_trace_blocks(aca_case, xavier_w):
    """Far blocks of the ACA test case with at most 160 x 160 particles, with
    their dense entries and the oracle's factors and stop-test trace."""
    x, m = aca_case
    e = m.export()
    tree, perm = e["tree"].astype(np.int64), e["perm"].astype(np.int64)
    out = []
    for b, (I, J) in enumerate(e["far"]):
        if tree[I, 3] - tree[I, 2] > 160 or tree[J, 3] - tree[J, 2] > 160:
            continue
        fa = m.aca_factors(xavier_w, b, 40)
        nu2, mu2 = m.aca_norms(xavier_w, b, 40)
        out.append((b, _sym_block(x, tree, perm, I, J, xavier_w), fa, nu2, mu2))
        if len(out) == 24:
            break
    assert len(out) >= 8
    return out


def test_aca_pivots_are_signed_determinant_argmax(aca_trace_blocks):
    """Every pivot is the first index of the largest SIGNED determinant of the
    residual 3x3 blocks over the unused indices (src/FarDot.cpp:537-582 for
    rows, :863-902 for columns), with the first column pivot |J|/2 (:257).
    Residuals are recomputed densely from the block and the returned
    factors.  The case is one where the max-|det| rule picks differently, so
    a |det| pivot fails here."""
    abs_differs = 0
    checked = 0
    for b, S, fa, _, _ in aca_trace_blocks:
        k, C, Q, ip, jp = fa["k"], fa["C"], fa["Q"], fa["ipiv"], fa["jpiv"]
        m, n = S.shape[:2]
        assert jp[0] == n // 2
        for it in range(k):
            # column residual at jpiv[it] and the row pivot
            col = S[:, jp[it]] - np.einsum("lirc,lcd->ird", C[:it], Q[:it, jp[it]])
            dets = np.linalg.det(col)
            elig = np.ones(m, dtype=bool)
            elig[ip[:it]] = False
            want, close = _pick(dets, elig)
            if not close:
                assert ip[it] == want, (b, it)
                checked += 1
            if int(np.argmax(np.where(elig, np.abs(dets), -np.inf))) != want:
                abs_differs += 1
            # the normalised factor is the residual column times the inverse
            # of its pivot block (:586-656): C_k times the pivot block gives
            # the residual back, to the rounding the inversion allows
            P = col[ip[it]]
            tol = 1e-12 * np.linalg.cond(P) * np.abs(C[it]).max() * np.abs(P).max()
            assert np.abs(C[it] @ P - col).max() < tol, (b, it)
            scale = np.abs(S).max()
            # row residual at ipiv[it] is the stored Q factor (:690-841)
            row = S[ip[it]] - np.einsum("lrc,ljcd->jrd", C[:it, ip[it]], Q[:it])
            assert np.abs(Q[it] - row).max() < 1e-10 * scale, (b, it)
            dq = np.linalg.det(Q[it])
            eq = np.ones(n, dtype=bool)
            eq[jp[:it + 1]] = False
            if eq.any():
                want, close = _pick(dq, eq)
                if not close:
                    assert jp[it + 1] == want, (b, it)
                    checked += 1
                if int(np.argmax(np.where(eq, np.abs(dq), -np.inf))) != want:
                    abs_differs += 1
    assert checked > 100
    assert abs_differs > 0, "case does not discriminate signed from |det| pivots"


def test_aca_nu2_is_frobenius_norm_of_new_term(aca_trace_blocks):
    """nu2 of iteration k is ||C_k Q_k||_F^2 over the whole block
    (src/FarDot.cpp:948-1037), from a dense product of the factors."""
    for b, S, fa, nu2, _ in aca_trace_blocks:
        for it in range(fa["k"]):
            term = np.einsum("ira,jab->ijrb", fa["C"][it], fa["Q"][it])
            np.testing.assert_allclose(nu2[it], np.sum(term ** 2), rtol=1e-9)


def _mu2_sequence(fa, nu2, last):
    """mu2 after each iteration when the cross terms of iteration k run over
    l < k - last (src/FarDot.cpp:1041-1078 reads l < k - 1)."""
    C, Q = fa["C"], fa["Q"]
    mu, out = 0.0, []
    for k in range(fa["k"]):
        mu += nu2[k]
        for l in range(k - last):
            cd = np.einsum("ira,irb->ab", C[k], C[l])
            qd = np.einsum("jac,jbc->ab", Q[k], Q[l])
            mu += 2.0 * np.sum(np.abs(cd) * np.abs(qd))
        out.append(mu)
    return np.array(out)


def test_aca_mu2_cross_terms_skip_previous_iteration(aca_trace_blocks):
    """The mu2 update adds |<C_k, C_l>| |<Q_k, Q_l>| cross terms for l < k-1
    only (src/FarDot.cpp:1041 loops to workingNodeIteration - 1 before the
    counter advances at :1094).  The rule is read from the reference, which
    cannot be run here; the test discriminates it from the textbook l < k
    sum, which the blocks below separate by more than rounding."""
    separated = 0
    for b, S, fa, nu2, mu2 in aca_trace_blocks:
        ref = _mu2_sequence(fa, nu2, 1)
        alt = _mu2_sequence(fa, nu2, 0)
        np.testing.assert_allclose(mu2, ref, rtol=1e-9)
        separated += int(np.any(np.abs(alt - ref) > 1e-6 * ref))
    assert separated > 0


def test_aca_stop_is_first_iteration_meeting_the_test(aca_trace_blocks):
    """The rank is the first iteration count with nu2 < mu2 eps^2, max_iter
    or min(|I|, |J|) reached (src/FarDot.cpp:1099-1108), eps = 1e-3."""
    stops = 0
    for b, S, fa, nu2, mu2 in aca_trace_blocks:
        m, n = S.shape[:2]
        k = fa["k"]
        met = [(nu2[i] < mu2[i] * 1e-6) or (i + 1 >= 40) or (i + 1 >= min(m, n)) for i in range(k)]
        assert met[-1] and not any(met[:-1]), b
        stops += int(nu2[k - 1] < mu2[k - 1] * 1e-6)
    assert stops > 0
