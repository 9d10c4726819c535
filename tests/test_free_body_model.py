"""Free-body models on the CPU: the scene description against newt's golden
bytes, closed forms of the oracle's integrators, and the model compiler's
rejections (no device calls: finalize validates before it allocates).
"""
import os

import numpy as np
import pytest

from tooling_amd import abi, scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _snapshot(q, v, nb):
    """(position, orientation, linear velocity, body angular velocity) per
    body as little-endian f32, newt/tests/solver_golden.rs:146-168."""
    out = []
    for b in range(nb):
        out += list(q[7 * b:7 * b + 7]) + list(v[6 * b:6 * b + 6])
    return np.asarray(out, dtype=np.float32).tobytes()


def test_box_stack_scene_is_newts_golden_stack(oracle_lib):
    """scenes.box_stack() through the C-ABI records, stepped by the oracle's
    float instantiation (newt's arithmetic), writes newt's solver_stack.bin
    byte for byte: the scene the GPU tests run is the reference's scene."""
    sc = scenes.box_stack()
    ow = oracle_lib.OracleWorlds(sc, 1, 1)
    expected = open(os.path.join(GOLDEN, "solver_stack.bin"), "rb").read()
    got = []
    q, v = ow.state()
    got.append(_snapshot(q[0], v[0], 3))
    done = 0
    for mark in (50, 200, 1000):
        ow.step(mark - done)
        done = mark
        q, v = ow.state()
        got.append(_snapshot(q[0], v[0], 3))
    ow.close()
    assert b"".join(got) == expected


def _golden_rollout(ow, nb, marks):
    q, v = ow.state()
    got = [_snapshot(q[0], v[0], nb)]
    done = 0
    for mark in marks:
        ow.step(mark - done)
        done = mark
        q, v = ow.state()
        got.append(_snapshot(q[0], v[0], nb))
    ow.close()
    return b"".join(got)


def test_penalty_stack_scene_is_newts_contacts_golden(oracle_lib):
    """scenes.penalty_stack() through the C-ABI records, stepped by the
    oracle's float instantiation under the penalty solver (spring-damper
    contact wrenches re-sampled at every RK4 stage, legacy vertex-face
    box-box manifold with the SAT fallback), writes newt's
    stacking_3_boxes.bin (newt/tests/contacts_golden.rs:82-97) byte for
    byte."""
    ow = oracle_lib.OracleWorlds(scenes.penalty_stack(), 1, 1)
    expected = open(os.path.join(GOLDEN, "stacking_3_boxes.bin"), "rb").read()
    assert _golden_rollout(ow, 3, (100, 1000)) == expected


def test_stack_json_penalty_is_newts_model_golden(oracle_lib):
    """newt's stack.json under its own (default penalty) solver, loaded by
    tooling_amd.model_json and stepped by the oracle's float instantiation,
    writes newt's model_stack.bin (newt/tests/model_load.rs:315-330) byte
    for byte."""
    from tooling_amd import model_json
    sc = model_json.load(os.path.join(GOLDEN, "models", "stack.json"))
    assert sc.options["solver"] == abi.SOLVER_PENALTY
    ow = oracle_lib.OracleWorlds(sc, 1, 1)
    expected = open(os.path.join(GOLDEN, "model_stack.bin"), "rb").read()
    assert _golden_rollout(ow, 3, (200, 500, 1000)) == expected


def test_free_fall_closed_form(oracle_lib):
    """Semi-implicit Euler with no contacts: v_n = v0 + n g dt and
    p_n = p0 + dt sum_{k=1..n} v_k = p0 + n dt v0 + g dt^2 n(n+1)/2; a body
    spinning about a principal axis keeps its angular velocity."""
    dt, n = 0.004, 250
    b = scenes.body(2.0, (0.1, 0, 0, 0, 0.2, 0, 0, 0, 0.3), (0.5, -1.0, 3.0), linear_velocity=(0.3, 0.0, 1.0),
                    angular_velocity_body=(0.0, 0.0, 2.0))
    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_PENALTY,
                iterations=1)
    sc = scenes.Scene("fall", opts, [], [], 0, None, bodies=[b])
    ow = oracle_lib.OracleWorlds(sc, 1, 0)
    ow.step(n)
    q, v = ow.state()
    ow.close()
    g = np.array([0.0, 0.0, -9.81])
    v0, p0 = np.array([0.3, 0.0, 1.0]), np.array([0.5, -1.0, 3.0])
    assert np.allclose(v[0, :3], v0 + n * g * dt, rtol=0, atol=1e-12)
    assert np.allclose(q[0, :3], p0 + n * dt * v0 + g * dt * dt * n * (n + 1) / 2, rtol=0, atol=1e-10)
    assert np.allclose(v[0, 3:], (0.0, 0.0, 2.0), atol=1e-15)
    # rotation about body z by n dt w
    ang = n * dt * 2.0
    assert np.allclose(q[0, 3:], (0.0, 0.0, np.sin(ang / 2), np.cos(ang / 2)), atol=1e-12)


def test_resting_box_carries_its_weight(oracle_lib):
    """A box settled on the plane: the PGS contact force balances gravity,
    so its velocity stays near zero and its height near the half extent."""
    sc = scenes.with_options(scenes.box_stack(n_boxes=1), "_euler", integrator=abi.INTEGRATOR_EULER)
    ow = oracle_lib.OracleWorlds(sc, 1, 0)
    ow.step(400)
    q, v = ow.state()
    a = ow.qacc()
    ow.close()
    assert abs(q[0, 2] - 0.25) < 2e-3
    assert np.abs(v[0, :3]).max() < 1e-3   # linear velocity (the PGS friction leaves a slow yaw drift)
    assert np.abs(a[0, :3]).max() < 0.05   # contact force cancels g within 0.5 %


def _library():
    try:
        return abi.load_library()
    except RuntimeError:
        pytest.skip("HIP library not built")


def _finalize(sc):
    from tooling_amd.newt import Model
    return Model(sc)


def _status(sc):
    _library()
    with pytest.raises(abi.NewtError) as e:
        _finalize(sc)
    return e.value.status


def test_newton_with_elliptic_cone_rejected():
    """SolverConfig::validate (solver.rs:123-131): Newton needs the pyramidal cone."""
    sc = scenes.with_options(scenes.box_stack(), "_newton_ell", solver=abi.SOLVER_NEWTON, cone=abi.CONE_ELLIPTIC)
    assert _status(sc) == abi.ERR_INVALID_CONFIG


def test_newton_stack_scene_is_newts_golden(oracle_lib):
    """scenes.box_stack() under the Newton solver, through the C-ABI records,
    writes newt's solver_newton_stack.bin byte for byte (oracle float)."""
    sc = scenes.with_options(scenes.box_stack(), "_newton", solver=abi.SOLVER_NEWTON)
    ow = oracle_lib.OracleWorlds(sc, 1, 1)
    expected = open(os.path.join(GOLDEN, "solver_newton_stack.bin"), "rb").read()
    q, v = ow.state()
    got = [_snapshot(q[0], v[0], 3)]
    done = 0
    for mark in (50, 200, 1000):
        ow.step(mark - done)
        done = mark
        q, v = ow.state()
        got.append(_snapshot(q[0], v[0], 3))
    ow.close()
    assert b"".join(got) == expected


def test_trees_with_bodies_unsupported_solvers_rejected():
    """A model with trees and free bodies compiles under PGS and Newton;
    penalty contacts are rejected at finalize before any device work."""
    st = scenes.box_stack()
    pend = scenes.pendulum()
    sc = scenes.Scene("mixed", dict(st.options, solver=abi.SOLVER_PENALTY), pend.links, st.geoms, 0, None,
                      bodies=st.bodies)
    assert _status(sc) == abi.ERR_UNSUPPORTED


def test_geom_on_missing_body_rejected():
    st = scenes.box_stack()
    geoms = st.geoms + [scenes.geom(abi.SHAPE_SPHERE, body=7, size=(0.1, 0, 0))]
    sc = scenes.Scene("bad", dict(st.options), [], geoms, 0, None, bodies=st.bodies)
    assert _status(sc) == abi.ERR_INVALID_ARG


def test_singular_inertia_rejected():
    st = scenes.box_stack()
    bodies = [dict(st.bodies[0], inertia=(1.0, 0, 0, 0, 0.0, 0, 0, 0, 1.0))] + st.bodies[1:]
    sc = scenes.Scene("bad", dict(st.options), [], st.geoms, 0, None, bodies=bodies)
    assert _status(sc) == abi.ERR_INVALID_ARG
