"""Pins of the oracle's free-body speed caps and gravity scale
(clamp_velocity, newt/src/world.rs:3273-3286; evaluate's gravity term,
world.rs:3223-3245) by newt's own stability tests
(newt/tests/stability.rs:35-114), on the float instantiation (newt's
arithmetic) with World::new's options: dt 5 ms, RK4, penalty solver
(world.rs:326-361).  The reference asserts exact equalities; so do these.
"""
import numpy as np
import pytest

from tooling_amd import abi, scenes

HALF = (0.5, 0.5, 0.5)


def _world(bodies, gravity=(0.0, 0.0, 0.0)):
    opts = dict(dt=0.005, gravity=gravity, integrator=abi.INTEGRATOR_RK4, solver=abi.SOLVER_PENALTY, iterations=20)
    return scenes.Scene("stability", opts, [], [], 0, None, bodies=bodies)


def _box(**kw):
    return scenes.body(1.0, scenes.solid_box_inertia(1.0, HALF), (0.0, 0.0, 0.0), **kw)


def _len32(v):
    v = np.asarray(v, dtype=np.float32)
    return np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2], dtype=np.float32)


def _run(oracle_lib, bodies, steps, gravity=(0.0, 0.0, 0.0)):
    ow = oracle_lib.OracleWorlds(_world(bodies, gravity), 1, 1)
    ow.step(steps)
    q, v = ow.state()
    ow.close()
    return q[0], v[0]


def test_max_linear_velocity(oracle_lib):
    _, v = _run(oracle_lib, [_box(linear_velocity=(10.0, 0.0, 0.0), max_linear_velocity=2.0)], 5)
    assert _len32(v[0:3]) == np.float32(2.0)
    _, v = _run(oracle_lib, [_box(linear_velocity=(10.0, 0.0, 0.0))], 5)
    assert _len32(v[0:3]) == np.float32(10.0)


def test_max_angular_velocity(oracle_lib):
    _, v = _run(oracle_lib, [_box(angular_velocity_body=(10.0, 0.0, 0.0), max_angular_velocity=2.0)], 5)
    assert _len32(v[3:6]) == np.float32(2.0)
    _, v = _run(oracle_lib, [_box(angular_velocity_body=(10.0, 0.0, 0.0))], 5)
    assert _len32(v[3:6]) == np.float32(10.0)


def test_gravity_scale(oracle_lib):
    q, _ = _run(oracle_lib, [_box(gravity_scale=0.0), _box()], 10, gravity=(0.0, -10.0, 0.0))
    assert q[1] == 0.0
    assert q[7 + 1] < -0.001


def test_gravity_scale_closed_form(oracle_lib):
    """A body with gravity scale s falls as s·g: after n RK4 steps of a
    constant acceleration, y = s g (n dt)^2 / 2 (RK4 is exact there)."""
    s, g, n, dt = 0.25, -10.0, 40, 0.005
    ow = oracle_lib.OracleWorlds(_world([_box(gravity_scale=s)], (0.0, g, 0.0)), 1, 0)
    ow.step(n)
    q, v = ow.state()
    ow.close()
    assert q[0, 1] == pytest.approx(0.5 * s * g * (n * dt) ** 2, rel=1e-12)
    assert v[0, 1] == pytest.approx(s * g * n * dt, rel=1e-12)
