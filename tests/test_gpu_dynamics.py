"""Device forcing and the device-resident integrators (GPU).

The forcing kernels are checked against oracle/forcing.py; the device
integrators against the host integrators of the pybind11 module driven by a
Python velocity callback that performs the same stage (forcing, update_coord,
dot) through the host API, so both sides see the same matvec kernels.
"""
import numpy as np
import pytest

from oracle import forcing as F

pytestmark = pytest.mark.gpu

f32 = np.float32


def hignn():
    from hignn_amd import hignn as h
    h.Init()
    return h


def suspension(seed, n_fil=8, n_chain=40, n_free=400, box=60.0, step=2.0):
    """n_fil random-walk chains (bond length `step`) then free particles."""
    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(n_fil):
        d = rng.normal(size=(n_chain, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        parts.append(rng.uniform(0.2 * box, 0.8 * box, 3) + np.cumsum(step * d, axis=0))
    parts.append(rng.uniform(0, box, size=(n_free, 3)))
    return np.concatenate(parts).astype(f32)


FIL = dict(n_fil=8, n_chain=40, rest=2.0, k_t=5.0, k_b=3.0)


def forcing(h, body=(0.0, 0.0, -1.0), morse=5.0):
    fc = h.DeviceForcing()
    fc.set_body_force(list(body))
    fc.set_filaments(FIL["n_fil"], FIL["n_chain"], FIL["rest"], FIL["k_t"], FIL["k_b"])
    if morse:
        fc.set_two_body_epsilon(morse)
    return fc


def test_device_forcing_matches_oracle():
    h = hignn()
    x = suspension(0)
    got = forcing(h).eval(x)
    want = F.velocity_force(x, (0, 0, -1), FIL["n_fil"], FIL["n_chain"], FIL["rest"], FIL["k_t"],
                            FIL["k_b"], morse=F.morse_force(x, 5.0))
    # same float32 terms on both sides; the sums differ only in order (double)
    scale = np.abs(want).max()
    err = np.abs(got - want).max() / scale
    print(f"forcing: max |f| {scale:.3g}, max rel diff {err:.2e}")
    assert err <= 1e-6


@pytest.mark.parametrize("eps", [3.0, 7.5])
def test_potential_force_matches_oracle(eps):
    h = hignn()
    rng = np.random.default_rng(1)
    x = rng.uniform(-20, 20, size=(1500, 3)).astype(f32)
    pf = h.PotentialForce()
    pf.set_two_body_epsilon(eps)
    got = pf.get_potential_force(x)
    want = F.morse_force(x, eps)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_potential_force_periodic_matches_oracle():
    h = hignn()
    rng = np.random.default_rng(2)
    lo, hi = np.array([-10, -12, 0], f32), np.array([14, 10, 25], f32)
    x = rng.uniform(lo, hi, size=(900, 3)).astype(f32)
    x[0] = lo + 0.5  # a corner particle: 8 images
    pf = h.PotentialForce()
    pf.set_two_body_epsilon(4.0)
    pf.set_domain(np.stack([lo, hi]))
    got = pf.get_potential_force(x)
    want = F.morse_force(x, 4.0, periodic=True, lo=lo, hi=hi)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    # set_periodic(False) returns to the plain search
    pf.set_periodic(False)
    np.testing.assert_allclose(pf.get_potential_force(x), F.morse_force(x, 4.0),
                               atol=1e-6 * np.abs(want).max())


def test_periodic_without_domain_is_rejected():
    h = hignn()
    pf = h.PotentialForce()
    pf.set_two_body_epsilon(2.0)
    pf.set_periodic(True)
    with pytest.raises(RuntimeError, match="domain"):
        pf.get_potential_force(np.zeros((4, 3), f32))


def model(h, x, w, reuse=False):
    m = h.HignnModel(x, 64)
    m.load_two_body_weights(w)
    m.set_epsilon(1e-3)
    m.set_max_iter(20)
    m.set_tree_reuse(reuse)
    return m


def integrate(h, cls, x, w, device, dt=0.02, tf=0.1, thr=None, lims=None, reuse=False,
              graphs=True, morse=5.0, stats=None, max_steps=None):
    m = model(h, x, w, reuse)
    fc = forcing(h, morse=morse)
    it = cls()
    it.set_time_step(dt)
    it.set_final_time(tf)
    it.set_num_rigid_body(x.shape[0])
    if thr is not None:
        it.set_threshold(thr)
    if lims is not None:
        it.set_x_lim(lims[0])
        it.set_y_lim(lims[1])
        it.set_z_lim(lims[2])
    outs = []
    it.set_output_func(lambda p, v: outs.append(np.array(p)))
    if device:
        it.set_device_velocity(m, fc)
        it.set_graphs(graphs)
        if max_steps is not None:
            it.set_max_steps(max_steps)
    else:
        def vel(t, pos):
            xx = np.ascontiguousarray(pos[:, :3])
            m.update_coord(xx)
            f = fc.eval(xx)
            u = np.zeros_like(f)
            m.dot(u, f)
            return u
        it.set_velocity_func(vel)
    state = np.zeros((x.shape[0], 6), f32)
    state[:, :3] = x
    it.initialize(state)
    it.run()
    if stats is not None:
        stats.append(it.device_stats())
    return it.positions(), it.func_count(), outs


def close(a, b, rel=1e-5):
    return np.abs(a - b).max() <= rel * np.abs(b).max()


@pytest.mark.parametrize("reuse", [False, True])
def test_device_euler_matches_host(xavier_w, reuse):
    h = hignn()
    x = suspension(3)
    xh, ch, oh = integrate(h, h.ExplicitEuler, x, xavier_w, device=False, reuse=reuse)
    xd, cd, od = integrate(h, h.ExplicitEuler, x, xavier_w, device=True, reuse=reuse)
    assert ch == cd == 5
    assert len(oh) == len(od) == 5
    print("euler max |x_dev - x_host|:", np.abs(xd - xh).max(), "moved:", np.abs(xh - x).max())
    assert np.abs(xh - x).max() > 1e-3  # the particles moved
    assert close(xd, xh)
    for a, b in zip(od, oh):
        assert close(a, b)


def test_device_euler_graphs_match_eager(xavier_w):
    h = hignn()
    x = suspension(4)
    xg, _, _ = integrate(h, h.ExplicitEuler, x, xavier_w, device=True, graphs=True)
    xe, _, _ = integrate(h, h.ExplicitEuler, x, xavier_w, device=True, graphs=False)
    assert close(xg, xe, 1e-6)


def test_device_euler_periodic_matches_host(xavier_w):
    h = hignn()
    x = suspension(5)
    lo, hi = x.min(axis=0) + 1.0, x.max(axis=0) - 1.0  # some particles start outside: wrapped
    lims = [[float(lo[j]), float(hi[j])] for j in range(3)]
    xh, _, _ = integrate(h, h.ExplicitEuler, x, xavier_w, device=False, lims=lims, morse=None)
    xd, _, _ = integrate(h, h.ExplicitEuler, x, xavier_w, device=True, lims=lims, morse=None)
    assert close(xd, xh)


def test_device_dopri_matches_host(xavier_w):
    h = hignn()
    x = suspension(6)
    # stiff bonds: the first step is rejected and shrunk, then a few accepted
    kw = dict(dt=0.002, tf=0.004, thr=1e-2)
    xh, ch, oh = integrate(h, h.ExplicitRk4, x, xavier_w, device=False, **kw)
    xd, cd, od = integrate(h, h.ExplicitRk4, x, xavier_w, device=True, **kw)
    assert ch == cd and len(oh) == len(od)
    print("dopri steps:", ch, "max |x_dev - x_host|:", np.abs(xd - xh).max())
    assert close(xd, xh)


@pytest.mark.parametrize("cls", ["ExplicitEuler", "ExplicitRk4"])
def test_device_max_steps_stops_at_the_full_runs_step(xavier_w, cls):
    """set_max_steps(k): the device-resident run stops after k accepted steps,
    at the positions the uncapped run reported after its k-th step."""
    h = hignn()
    x = suspension(6)
    kw = dict(dt=0.002, tf=0.0006, thr=1e-2) if cls == "ExplicitRk4" else dict(dt=0.02, tf=0.1)
    c = getattr(h, cls)
    xf, cf, of = integrate(h, c, x, xavier_w, device=True, **kw)
    assert cf >= 2 and len(of) == cf
    k = cf - 1
    xk, ck, ok = integrate(h, c, x, xavier_w, device=True, max_steps=k, **kw)
    print(f"{cls}: {cf} steps uncapped, stopped after {ck}")
    assert ck == k and len(ok) == k
    assert close(xk, of[k - 1], 1e-6)
    with pytest.raises(RuntimeError):
        c().set_max_steps(-1)


def test_device_euler_with_three_body_matches_host(xavier_w):
    # the velocity of python/gravity_field.py:58-75: matvec + three-body and self corrections
    from hignn_amd import weights as W
    h = hignn()
    x = suspension(7, n_fil=4, n_chain=20, n_free=200)
    tb_w = (W.xavier(1, W.net_shapes(6)), W.xavier(2, W.net_shapes(3)))

    def run(device):
        m = model(h, x, xavier_w)
        fc = h.DeviceForcing()
        fc.set_body_force([0.0, 0.0, -1.0])
        tb = h.ThreeBodyModel()
        tb.load_weights(*tb_w)
        tb.set_three_body_epsilon(5.0)
        it = h.ExplicitEuler()
        it.set_time_step(0.01)
        it.set_final_time(0.03)
        it.set_num_rigid_body(x.shape[0])
        if device:
            it.set_device_velocity(m, fc, tb)
        else:
            def vel(t, pos):
                xx = np.ascontiguousarray(pos[:, :3])
                m.update_coord(xx)
                f = fc.eval(xx)
                u = np.zeros_like(f)
                m.dot(u, f)
                return u + tb.velocity(xx, f)
            it.set_velocity_func(vel)
        state = np.zeros((x.shape[0], 6), f32)
        state[:, :3] = x
        it.initialize(state)
        it.run()
        return it.positions()

    xh, xd = run(False), run(True)
    print("euler + three-body max |x_dev - x_host|:", np.abs(xd - xh).max())
    assert np.abs(xh - x).max() > 1e-4
    assert close(xd, xh)


def test_forcing_edge_cases(xavier_w):
    h = hignn()
    x = suspension(8)
    fc = h.DeviceForcing()
    assert not fc.eval(x).any()  # no term set: zero force
    fc.set_filaments(100, 40, 2.0, 1.0, 1.0)  # 4000 beads > N
    with pytest.raises(RuntimeError, match="filaments"):
        fc.eval(x)
    fc.set_filaments(0, 0, 0.0, 0.0, 0.0)
    fc.set_two_body_epsilon(5.0)
    fc.set_domain(np.array([[0, 0, 0], [8, 100, 100]], f32))  # 8 < 2 x 5
    with pytest.raises(RuntimeError, match="twice the radius"):
        fc.eval(x)
    # integrator and model of different sizes
    it = h.ExplicitEuler()
    it.set_num_rigid_body(x.shape[0] + 1)
    with pytest.raises(RuntimeError, match="sizes differ"):
        it.set_device_velocity(model(h, x, xavier_w), h.DeviceForcing())


def test_stepper_recaptures_after_forcing_buffers_move(xavier_w):
    # the stage graphs hold the forcing's cell-list buffers; an evaluation of a
    # larger set through the same forcing moves them, and the stepper must
    # capture its graphs again before the next run
    h = hignn()
    x = suspension(9, n_fil=2, n_chain=20, n_free=160)
    m = model(h, x, xavier_w)
    fc = h.DeviceForcing()
    fc.set_body_force([0.0, 0.0, -1.0])
    fc.set_filaments(2, 20, FIL["rest"], FIL["k_t"], FIL["k_b"])
    fc.set_two_body_epsilon(5.0)
    it = h.ExplicitEuler()
    it.set_time_step(0.02)
    it.set_final_time(0.04)
    it.set_num_rigid_body(x.shape[0])
    it.set_device_velocity(m, fc)
    state = np.zeros((x.shape[0], 6), f32)
    state[:, :3] = x

    def run():
        it.initialize(state)
        it.run()
        return it.positions()

    first = run()
    fc.eval(suspension(10, n_fil=2, n_chain=20, n_free=4000))  # grows the cell list
    assert close(run(), first, 1e-6)


@pytest.mark.parametrize("method", ["euler", "dopri"])
def test_whole_stage_graphs_match_eager(xavier_w, monkeypatch, method):
    # tree reuse and the persistent far-field kernel: after the first (eager)
    # evaluation every stage (stage input, forcing, coordinate gather, matvec)
    # replays as one graph.  Every step's positions are checked against the
    # eager stepper; the output callback synchronises with the host between
    # replays.
    monkeypatch.setenv("HH_FAR_MODE", "0")
    h = hignn()
    x = suspension(11)
    cls, kw = ((h.ExplicitEuler, dict(dt=0.02, tf=0.12)) if method == "euler" else
               (h.ExplicitRk4, dict(dt=0.002, tf=0.006, thr=1e-2)))
    sg, se = [], []
    xg, cg, og = integrate(h, cls, x, xavier_w, device=True, reuse=True, graphs=True, stats=sg, **kw)
    xe, ce, oe = integrate(h, cls, x, xavier_w, device=True, reuse=True, graphs=False, stats=se,
                           **kw)
    print(method, "stats graphs:", sg[0], "eager:", se[0])
    assert cg == ce and len(og) == len(oe) >= 3
    for a, b in zip(og, oe):
        assert close(a, b, 1e-6)
    assert close(xg, xe, 1e-6)
    assert np.abs(xg - x).max() > 1e-3
    assert sg[0]["whole_stage_graphs"] == sg[0]["velocity_evaluations"] - 1
    assert se[0]["whole_stage_graphs"] == 0


def test_whole_stage_graphs_recapture_after_model_settings(xavier_w, monkeypatch):
    # a new ACA tolerance and iteration cap change what the captured matvec
    # launches with: the stepper captures again and still matches the eager one
    monkeypatch.setenv("HH_FAR_MODE", "0")
    h = hignn()
    x = suspension(12)
    state = np.zeros((x.shape[0], 6), f32)
    state[:, :3] = x

    def make(graphs):
        m = model(h, x, xavier_w, reuse=True)
        it = h.ExplicitEuler()
        it.set_time_step(0.02)
        it.set_final_time(0.08)
        it.set_num_rigid_body(x.shape[0])
        it.set_device_velocity(m, forcing(h))
        it.set_graphs(graphs)
        return m, it

    def run(it):
        it.initialize(state)
        it.run()
        return it.positions(), it.device_stats()

    (mg, itg), (me, ite) = make(True), make(False)
    a, sa = run(itg)
    b, _ = run(ite)
    assert close(a, b, 1e-6) and sa["whole_stage_graphs"] == sa["velocity_evaluations"] - 1
    for m in (mg, me):
        m.set_epsilon(1e-2)
        m.set_max_iter(8)
    c, sc = run(itg)
    d, _ = run(ite)
    print("before:", sa, "after:", sc)
    assert close(c, d, 1e-6)
    # one eager evaluation sets the persistent kernel up for the new cap
    assert sc["whole_stage_graphs"] == sc["velocity_evaluations"] - 1


def test_tree_rebuild_every_step_equals_rebuilding_stepper(xavier_w):
    """With the tree kept across stages but rebuilt at every accepted step,
    an Euler run (one stage per step) builds the same trees from the same
    positions as the stepper without tree reuse: identical trajectories."""
    h = hignn()
    x = suspension(3)
    kw = dict(dt=0.02, tf=0.1)
    xa, ca, _ = integrate(h, h.ExplicitEuler, x, xavier_w, device=True, **kw)
    m = model(h, x, xavier_w, reuse=True)
    it = h.ExplicitEuler()
    it.set_time_step(kw["dt"])
    it.set_final_time(kw["tf"])
    it.set_num_rigid_body(x.shape[0])
    it.set_device_velocity(m, forcing(h))
    it.set_tree_rebuild(1)
    state = np.zeros((x.shape[0], 6), f32)
    state[:, :3] = x
    it.initialize(state)
    it.run()
    assert it.func_count() == ca
    np.testing.assert_array_equal(it.positions(), xa)


def test_rebuild_after_device_updates_reads_current_positions(xavier_w):
    """Device-side coordinate updates on a kept tree leave the host copy
    behind; a later rebuild (here forced by a settings change) must build
    from the current positions."""
    from hignn_amd import abi, workloads
    x = workloads.uniform_cube(900, 0.05, 4)
    x2 = (x + np.random.default_rng(4).normal(scale=3.0, size=x.shape)).astype(f32)
    m = abi.Model(x, 64)
    m.load_weights(xavier_w)
    m.set(epsilon=1e-3, max_iter=20, tree_reuse=True)
    m.update_coord()
    xd = abi.DeviceArray.of(x2)
    m.update_coord_device(xd.ptr)
    m.set(use_symmetry=False)  # clears the built tree: the next dot rebuilds
    f = np.random.default_rng(5).normal(size=x.shape).astype(f32)
    m.dot(f)
    fresh = abi.Model(x2, 64)
    fresh.load_weights(xavier_w)
    fresh.set(epsilon=1e-3, max_iter=20, use_symmetry=False)
    fresh.update_coord()
    a, b = m.export(), fresh.export()
    for k in ("perm", "tree", "close", "far"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
