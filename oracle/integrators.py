"""numpy restatement of the reference time integrators (TEST INFRASTRUCTURE).

ExplicitEuler::Run (include/TimeIntegrator.hpp:431-488) and ExplicitRk4::Run
(include/TimeIntegrator.hpp:602-860), positions as float32 triples with the
reference's operation order, no periodic wrap, velocities from ``vel(t, x)``.
Used only by tests/ to check the integrators of the pybind11 module.
"""
import numpy as np

f32 = np.float32


def euler(x0, vel, dt, tf):
    x = np.array(x0, dtype=f32)
    # set_time_step / set_final_time take float (TimeIntegrator.hpp:271-281)
    dt, tf = float(f32(dt)), float(f32(tf))
    t, h0, calls = 0.0, float(dt), 0
    while t < tf - 1e-5:
        d = min(h0, tf - t)
        v = np.asarray(vel(f32(t + d), x), dtype=f32)
        x = (x + f32(d) * v).astype(f32)
        t += d
        calls += 1
    return x, calls


# Dormand-Prince 5(4) tableau in float32 (include/TimeIntegrator.hpp:504-544)
A = [[f32(1) / f32(5)],
     [f32(3) / f32(40), f32(9) / f32(40)],
     [f32(44) / f32(45), f32(-56) / f32(15), f32(32) / f32(9)],
     [f32(19372) / f32(6561), f32(-25360) / f32(2187), f32(64448) / f32(6561), f32(-212) / f32(729)],
     [f32(9017) / f32(3168), f32(-355) / f32(33), f32(46732) / f32(5247), f32(49) / f32(176),
      f32(-5103) / f32(18656)]]
B = [f32(35) / f32(384), f32(0), f32(500) / f32(1113), f32(125) / f32(192),
     f32(-2187) / f32(6784), f32(11) / f32(84)]
DC = [B[0] - f32(5179) / f32(57600), f32(0), B[2] - f32(7571) / f32(16695),
      B[3] - f32(393) / f32(640), B[4] - f32(-92097) / f32(339200), B[5] - f32(187) / f32(2100),
      f32(-1) / f32(40)]
C = [f32(0), f32(1) / f32(5), f32(3) / f32(10), f32(4) / f32(5), f32(8) / f32(9), f32(1), f32(1)]


def dopri(x0, vel, dt, tf, thr=1e-3):
    """Returns final positions, number of accepted steps, the accepted step
    sizes and the number of velocity calls."""
    x = np.array(x0, dtype=f32)
    dt, tf = float(f32(dt)), float(f32(tf))
    n = x.shape[0]
    K = [np.zeros((n, 3), dtype=f32) for _ in range(7)]  # K1 starts at zero
    t, h0 = 0.0, float(dt)
    calls = 1
    vel(f32(t), x)  # the first call's result is not stored in K1
    steps, sizes = 0, []
    while t < tf - 1e-5:
        d = min(h0, tf - t)
        x0s = x.copy()
        err = 100.0
        while err > thr:
            for s in range(1, 7):
                xs = x0s.copy()
                coefs = A[s - 1] if s < 6 else B
                for k, a in enumerate(coefs):
                    if s == 6 and k == 1:
                        continue
                    xs = (xs + f32(float(a) * d) * K[k]).astype(f32)
                x = xs
                K[s] = np.asarray(vel(f32(t + float(C[s]) * d), x), dtype=f32)
                calls += 1
            v = (K[0] * DC[0] + K[2] * DC[2]).astype(f32)
            for k in (3, 4, 5, 6):
                v = (v + K[k] * DC[k]).astype(f32)
            v = (f32(d) * v).astype(f32)
            e2 = ((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]).astype(f32) + v[:, 2] * v[:, 2]).astype(f32)
            err = float(np.sqrt(np.sum(e2.astype(np.float64))))
            if err > thr:
                d = d * max(0.8 * (err / thr) ** -0.2, 0.1)
                d = max(d, 1e-6)
        t += d
        sizes.append(d)
        K[0] = K[6]
        steps += 1
    return x, steps, sizes, calls
