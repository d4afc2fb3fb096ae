"""Many independent suspensions in one model (hh_model_create_instances)
against one model per suspension, and per-instance errors of the chosen form
against the FP64 evaluation (GPU)."""
import numpy as np
import pytest

from hignn_amd import abi, weights, workloads
from oracle import surrogate as S

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def case():
    x, ns = workloads.instances(count=6, nmin=300, nmax=3000, seed=7)
    f = np.random.default_rng(7).normal(size=x.shape).astype(np.float32)
    return x, ns, f, weights.smooth(0)


def setup(m, w, prec="fp32"):
    m.load_weights(w)
    m.set(epsilon=1e-4, max_iter=30, precision=prec)
    m.update_coord()
    return m


def test_instances_equal_separate_models(case):
    x, ns, f, w = case
    mb = setup(abi.Model(x, 64, instances=ns), w)
    ub = mb.dot(f)
    eb = mb.export()
    rb = mb.far_ranks()
    off = np.concatenate([[0], np.cumsum(ns)])
    nodes, fars, closes = 0, 0, 0
    for k in range(len(ns)):
        sl = slice(off[k], off[k + 1])
        m = setup(abi.Model(x[sl], 64), w)
        u = m.dot(f[sl])
        e = m.export()
        # the merged tree and lists are the instance's own, shifted
        T, nc, nf = len(e["tree"]), len(e["close"]), len(e["far"])
        tb = eb["tree"][nodes:nodes + T].astype(np.int64)
        te = e["tree"].astype(np.int64)
        leaf = te[:, 0] == 0
        np.testing.assert_array_equal(tb[:, 0], np.where(leaf, 0, te[:, 0] + nodes))
        np.testing.assert_array_equal(tb[:, 2:4], te[:, 2:4] + off[k])
        np.testing.assert_array_equal(eb["perm"][off[k]:off[k + 1]], e["perm"] + off[k])
        np.testing.assert_array_equal(eb["close"][closes:closes + nc], e["close"] + nodes)
        np.testing.assert_array_equal(eb["far"][fars:fars + nf], e["far"] + nodes)
        np.testing.assert_array_equal(rb[fars:fars + nf], m.far_ranks())
        # same blocks and ranks: the velocities differ only in atomic order
        assert rel(ub[sl], u) <= 1e-6, k
        nodes, fars, closes = nodes + T, fars + nf, closes + nc
    assert nodes == len(eb["tree"]) and fars == len(eb["far"]) and closes == len(eb["close"])


def test_instances_chosen_form_per_instance_error(case):
    # gravity, as the many-instance workload; with random forces the sums
    # cancel and the BF16 near-field rounding (about 1e-3 per block entry)
    # is no longer averaged down
    x, ns, _, w = case
    f = workloads.gravity(x.shape[0])
    m0 = abi.Model(workloads.uniform_cube(200, 0.05, 0), 64)
    m0.load_weights(w)
    n = 1 << 21
    r, t = m0.surrogate_samples(n, 2.0, 650.0, seed=1)
    s = abi.Surrogate(2.0, 650.0)
    s.set_params(S.init(0))
    s.train(r.ptr, t.ptr, n, 20000, 3e-3, batch=16384, lr_final=3e-5)
    mb = setup(abi.Model(x, 64, instances=ns), w, prec="bf16")
    mb.set_surrogate(s)
    u = mb.dot(f)
    off = np.concatenate([[0], np.cumsum(ns)])
    tg = np.concatenate([np.arange(off[k], off[k + 1], max(1, ns[k] // 100)) for k in range(len(ns))])
    ref = mb.check_fp64(f, tg)
    errs = []
    for k in range(len(ns)):
        sel = (tg >= off[k]) & (tg < off[k + 1])
        errs.append(rel(u[tg[sel]], ref[sel]))
    st = mb.surrogate_stats()
    print(f"per-instance error vs FP64: {['%.1e' % e for e in errs]}; surrogate blocks {st[1]} of {st[0]}")
    assert max(errs) <= 1e-3
