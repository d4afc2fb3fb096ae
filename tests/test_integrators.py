"""ExplicitEuler / ExplicitRk4 of the pybind11 module (no device use)
against closed forms and the numpy restatement in oracle/integrators.py."""
import numpy as np
import pytest

from oracle import integrators as R


def hignn():
    from hignn_amd import hignn as h
    return h


def run(cls, x0, vel, dt, tf, thr=None, out=None):
    it = cls()
    it.set_time_step(dt)
    it.set_final_time(tf)
    it.set_num_rigid_body(x0.shape[0])
    if thr is not None:
        it.set_threshold(thr)
    calls = []

    def cb(t, pos):
        calls.append((float(t), np.array(pos)))
        return vel(t, np.ascontiguousarray(pos[:, :3]))

    it.set_velocity_func(cb)
    got = []
    it.set_output_func(lambda p, v: got.append((np.array(p), np.array(v))))
    if out:
        it.set_output_step(out)
    state = np.zeros((x0.shape[0], 6), np.float32)
    state[:, :3] = x0
    it.initialize(state)
    it.run()
    return got, calls


def test_euler_constant_velocity_exact():
    x0 = np.random.default_rng(0).normal(size=(20, 3)).astype(np.float32)
    c = np.array([0.5, -1.0, 0.25], np.float32)
    got, calls = run(hignn().ExplicitEuler, x0, lambda t, x: np.tile(c, (x.shape[0], 1)), 0.125, 1.0)
    assert len(calls) == 8
    np.testing.assert_allclose(got[-1][0], x0 + c, rtol=0, atol=2e-6)
    # callback receives [x, y, z, roll, pitch, yaw] (N, 6)
    assert calls[0][1].shape == (20, 6)


def test_euler_matches_restatement():
    g = np.random.default_rng(1)
    x0 = g.normal(size=(50, 3)).astype(np.float32)
    A = g.normal(size=(3, 3)).astype(np.float32) * 0.3
    vel = lambda t, x: (x @ A.T).astype(np.float32)
    got, _ = run(hignn().ExplicitEuler, x0, vel, 0.01, 0.1)
    want, _ = R.euler(x0, vel, 0.01, 0.1)
    np.testing.assert_array_equal(got[-1][0], want)


def test_rk4_first_step_quirk_closed_form():
    """K1 is zero on the first step (include/TimeIntegrator.hpp:608), so a
    constant field moves the first accepted step by (1 - b1) dt c; from the
    second step on (K1 = K7 = c) the step is exact."""
    x0 = np.zeros((1, 3), np.float32)
    c = np.array([1.0, 0.0, 0.0], np.float32)
    vel = lambda t, x: np.tile(c, (x.shape[0], 1))
    want, steps, sizes, _ = R.dopri(x0, vel, 0.01, 0.02, thr=1.0)
    assert steps == 2
    b1 = 35.0 / 384.0
    assert abs(float(want[0, 0]) - (0.02 - b1 * 0.01)) < 1e-6
    got, _ = run(hignn().ExplicitRk4, x0, vel, 0.01, 0.02, thr=1.0)
    np.testing.assert_array_equal(got[-1][0], want)


def test_rk4_step_rejection_matches_restatement():
    """A tight threshold forces rejected steps (dt shrinks by
    max(0.8 (err/thr)^-0.2, 0.1)); the step size resets every step."""
    g = np.random.default_rng(2)
    x0 = g.normal(size=(30, 3)).astype(np.float32)
    A = g.normal(size=(3, 3)).astype(np.float32)
    vel = lambda t, x: (x @ A.T).astype(np.float32)
    want, steps, sizes, ncalls = R.dopri(x0, vel, 0.05, 0.2, thr=1e-4)
    assert min(sizes) < 0.05  # some rejection happened
    got, calls = run(hignn().ExplicitRk4, x0, vel, 0.05, 0.2, thr=1e-4)
    assert len(calls) == ncalls
    np.testing.assert_array_equal(got[-1][0], want)


def test_rk4_linear_field_accuracy():
    """From the second step on DOPRI is 5th order: a slow linear field is
    followed to the exact exp(At) x0 closely."""
    from scipy.linalg import expm
    g = np.random.default_rng(3)
    x0 = g.normal(size=(10, 3)).astype(np.float32)
    A = (g.normal(size=(3, 3)) * 0.2).astype(np.float32)
    vel = lambda t, x: (x @ A.T).astype(np.float32)
    got, _ = run(hignn().ExplicitRk4, x0, vel, 0.01, 0.5, thr=1e-2)
    exact = x0.astype(np.float64) @ expm(0.5 * A.astype(np.float64)).T
    # the first step loses b1*dt*K1; later steps are accurate
    err = np.abs(got[-1][0] - exact).max() / np.abs(exact).max()
    assert err < 5e-3


def test_init_shape_errors():
    it = hignn().ExplicitEuler()
    it.set_num_rigid_body(4)
    with pytest.raises(RuntimeError):
        it.initialize(np.zeros((3, 6), np.float32))
    with pytest.raises(RuntimeError):
        it.initialize(np.zeros(24, np.float32))
