"""The allocation-free surrogate gradient (hh_surrogate_gradient_noalloc)
against the allocating entry point, through the C ABI (GPU)."""
import numpy as np
import pytest

from hignn_amd import abi
from oracle import surrogate as S

pytestmark = pytest.mark.gpu


def case(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = (d * np.exp(rng.uniform(np.log(4), np.log(200), size=(n, 1)))).astype(np.float32)
    t = rng.normal(size=(n, 6)).astype(np.float32)
    return r, t


def test_noalloc_gradient_equals_allocating_gradient():
    s = abi.Surrogate(4.0, 200.0)
    p = (S.init(5) + np.random.default_rng(6).normal(size=S.NPARAMS) * 0.05).astype(np.float32)
    s.set_params(p)
    grad = abi.DeviceArray((S.NPARAMS,))
    loss = abi.DeviceArray((1,))
    # two calls in a row, the second smaller: the workspace keeps the first
    # call's partials, which the second must not pick up; the first is large
    # enough to use every block of the workspace
    for n, seed in ((200000, 1), (777, 2)):
        r, t = case(n, seed)
        rd, td = abi.DeviceArray.of(r), abi.DeviceArray.of(t)
        l_ref, g_ref = s.gradient(rd.ptr, td.ptr, n)
        grad.upload(np.full(S.NPARAMS, np.nan, np.float32))
        loss.upload(np.full(1, np.nan, np.float32))
        s.gradient_noalloc(rd.ptr, td.ptr, n, grad.ptr, loss.ptr)
        abi.synchronize()
        g, l = grad.numpy(), float(loss.numpy()[0])
        np.testing.assert_array_equal(g, g_ref)
        assert l == l_ref
        _, g64 = S.loss_grad(p, r, t, 200.0)
        assert np.linalg.norm(g - g64) <= 1e-5 * np.linalg.norm(g64)


def test_noalloc_gradient_rejects_bad_arguments():
    s = abi.Surrogate(4.0, 200.0)
    grad = abi.DeviceArray((S.NPARAMS,))
    loss = abi.DeviceArray((1,))
    r, t = case(10, 3)
    rd, td = abi.DeviceArray.of(r), abi.DeviceArray.of(t)
    with pytest.raises(RuntimeError, match="sample count"):
        s.gradient_noalloc(rd.ptr, td.ptr, 0, grad.ptr, loss.ptr)
    with pytest.raises(RuntimeError, match="NULL"):
        s.gradient_noalloc(rd.ptr, td.ptr, 10, 0, loss.ptr)
