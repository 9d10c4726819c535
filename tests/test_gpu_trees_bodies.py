"""Trees and free bodies in one model on the GPU (World::step with both,
newt/src/world.rs:1459-1537): the contacts touching a link are solved in the
tree solve, with a free body on the other side entering its rows as a body
component (newt/src/solver.rs:2805-3366, 3537-3836); the rest go to the
free-body solve, whose wrenches take the tree solve's body wrenches on top
(world.rs:2097-2149).  Such models run the lane-per-world FP64 step in every
precision, so each case is bitwise with the FP64 oracle, whose tree-body
rows newt's anchors and a reduction to the tree-tree solve pin
(test_oracle_tree_body.py):

- newt's sphere tree and sphere body pushed apart, and its tilted tree box
  on a box body, under PGS and Newton;
- a hinged box chain on a plane with free boxes dropped on it, under PGS
  with Euler and RK4 and under Newton with Euler: tree-body, body-body,
  body-plane and tree-plane contacts, contact lists equal every 10th step,
  state and qacc bitwise after 300 steps;
- newt's pendulum model with a ball, loaded from JSON;
- state sensors on links and bodies, bitwise with the oracle;
- rejections: penalty contacts, acceleration/force/touch sensors.
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu

PRECISIONS = [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64, abi.PRECISION_MIXED]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _kind(g):
    return "tree" if g["tree"] >= 0 else "body" if g["body"] >= 0 else "static"


def _run(oracle_lib, sc, n, steps, precision, seed=3, every=10, perturb=0.0):
    q0, v0 = sc.initial_state(n, seed)
    if perturb:
        v0 = v0 + np.random.default_rng(seed).normal(0.0, perturb, size=v0.shape)
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle_lib.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    kinds = {}
    for s in range(steps):
        step(m, d, 1)
        ow.step(1, threads=8)
        if s % every == every - 1 or s == steps - 1:
            counts, gids, _ = d.contacts()
            for w in range(n):
                oids, _ = ow.contacts(w)
                got = [tuple(x) for x in gids[w, :counts[w]]]
                assert got == [tuple(x) for x in oids], (s, w, got, oids)
                for a, b, _slot in got:
                    k = tuple(sorted((_kind(sc.geoms[a]), _kind(sc.geoms[b]))))
                    kinds[k] = kinds.get(k, 0) + 1
    assert not d.faults().any(), np.unique(d.faults())
    q, v = d.state()
    oq, ov = ow.state()
    np.testing.assert_array_equal(q, oq)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(d.qacc(), ow.qacc())
    d.close()
    ow.close()
    return kinds


@pytest.mark.parametrize("solver", [abi.SOLVER_PGS, abi.SOLVER_NEWTON])
def test_tree_body_anchor_scenes(oracle_lib, solver):
    for sc in (scenes.tree_body_spheres(solver), scenes.tree_body_tilted_boxes(solver),
               scenes.tree_body_tilted_boxes(solver, abi.INTEGRATOR_EULER)):
        for precision in PRECISIONS:
            _say(f"{sc.name} solver {solver} precision {precision}")
            kinds = _run(oracle_lib, sc, 16, 60, precision, every=1, perturb=0.05)
            assert kinds.get(("body", "tree"), 0) > 0


@pytest.mark.parametrize("solver,integrator", [(abi.SOLVER_PGS, abi.INTEGRATOR_EULER),
                                              (abi.SOLVER_PGS, abi.INTEGRATOR_RK4),
                                              (abi.SOLVER_NEWTON, abi.INTEGRATOR_EULER)])
@pytest.mark.parametrize("precision", PRECISIONS)
def test_chain_with_free_boxes(oracle_lib, solver, integrator, precision):
    sc = scenes.chain_and_boxes(solver=solver, integrator=integrator)
    _say(f"{sc.name} solver {solver} integrator {integrator} precision {precision}: 32 worlds x 300 steps")
    kinds = _run(oracle_lib, sc, 32, 300, precision)
    _say(f"contact kinds {kinds}")
    for k in (("body", "tree"), ("body", "body"), ("body", "static"), ("static", "tree")):
        assert kinds.get(k, 0) > 0, (k, kinds)


def test_trees_with_bodies_rejections():
    base = scenes.chain_and_boxes()
    sc = scenes.with_options(base, "_x", solver=abi.SOLVER_PENALTY)
    with pytest.raises(abi.NewtError) as e:
        Model(sc)
    assert e.value.status == abi.ERR_UNSUPPORTED


def _pendulum_with_ball():
    """newt's pendulum model (newt/models/pendulum.json) with a sphere on its
    lower link, a plane, and a free ball dropped on the lower link."""
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "models", "pendulum.json")) as f:
        doc = json.load(f)
    doc.pop("sites", None)
    doc["integrator"] = "Euler"
    doc["timestep"] = 0.002
    doc["solver"] = {"mode": "pgs", "iterations": 30}
    doc["trees"][0]["links"][2]["joint"]["axis"] = [1.0, 0.0, 0.0]
    doc["bodies"] = [{"name": "ball", "mass": 0.4, "inertia": {"kind": "diag", "values": [0.0016, 0.0016, 0.0016]},
                      "pose": {"position": [0.0, 0.03, 0.95]}}]
    doc["geoms"] = [
        {"shape": {"kind": "plane"}, "friction": 0.8},
        {"shape": {"kind": "sphere", "radius": 0.1}, "attach": {"kind": "link", "tree": "pendulum", "link": "lower"},
         "local_offset": [0.0, 0.0, 0.0], "friction": 0.8},
        {"shape": {"kind": "sphere", "radius": 0.1}, "attach": {"kind": "body", "body": "ball"}, "friction": 0.8},
    ]
    return json.dumps(doc)


def test_trees_with_bodies_from_json(oracle_lib):
    """A JSON model with a tree and a free body (model.rs:179-195): both
    loaders accept it; the GPU steps it bitwise with the oracle from the
    model's default state, the ball landing on the pendulum."""
    import ctypes as C
    from tooling_amd import model_json
    text = _pendulum_with_ball()
    lib = abi.load_library()
    h = C.c_void_p()
    abi.check(lib.newt_model_from_json(text.encode(), 1, C.byref(h)))
    nt, nb, ng = C.c_int32(), C.c_int32(), C.c_int32()
    abi.check(lib.newt_model_counts(h, C.byref(nt), C.byref(nb), C.byref(ng)))
    assert (nt.value, nb.value, ng.value) == (1, 1, 3)
    lib.newt_model_destroy(h)
    sc = model_json.load(text)
    m = Model(sc)
    d = Data(m, 4, abi.PRECISION_FP64)
    ow = oracle_lib.OracleWorlds(sc, 4, 0)
    q_gpu, _ = d.state()
    q_or, _ = ow.state()
    np.testing.assert_array_equal(q_gpu, q_or)
    tree_body = 0
    for s in range(400):
        step(m, d, 1)
        ow.step(1, threads=4)
        counts, gids, _ = d.contacts()
        for a, b, _slot in gids[0, :counts[0]]:
            if {_kind(sc.geoms[a]), _kind(sc.geoms[b])} == {"tree", "body"}:
                tree_body += 1
    assert not d.faults().any()
    q, v = d.state()
    oq, ov = ow.state()
    np.testing.assert_array_equal(q, oq)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(d.qacc(), ow.qacc())
    _say(f"pendulum with ball: {tree_body} tree-body contact-steps")
    assert tree_body > 0
    d.close()
    ow.close()


@pytest.mark.parametrize("precision", PRECISIONS)
def test_sensors_on_trees_with_bodies(oracle_lib, precision):
    """State sensors on links and on free bodies of one model (sensor.rs:
    446-702, evaluated after the step): joint position, frame position and
    orientation, gyro, velocimeter, frame linear and angular velocity with
    site offsets; bitwise with the oracle after 150 steps.  Acceleration,
    force and touch sensors on such models are rejected at finalize."""
    S = scenes.sensor
    sc = scenes.chain_and_boxes()
    sc.sensors = [S(abi.SENSOR_JOINT_POS, tree=0, link=1),
                  S(abi.SENSOR_FRAME_POS, tree=0, link=2, offset=(0.05, 0.0, 0.01)),
                  S(abi.SENSOR_FRAME_QUAT, tree=0, link=3),
                  S(abi.SENSOR_GYRO, body=0),
                  S(abi.SENSOR_VELOCIMETER, body=1, offset=(0.02, -0.01, 0.0)),
                  S(abi.SENSOR_FRAME_LINVEL, body=2, offset=(0.0, 0.03, 0.0)),
                  S(abi.SENSOR_FRAME_ANGVEL, tree=0, link=0)]
    n = 16
    q0, v0 = sc.initial_state(n, 5)
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle_lib.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    step(m, d, 150)
    ow.step(150, threads=8)
    assert not d.faults().any()
    np.testing.assert_array_equal(d.state()[0], ow.state()[0])
    np.testing.assert_array_equal(d.sensordata(), ow.sensordata())
    d.close()
    ow.close()
    bad = scenes.chain_and_boxes()
    bad.sensors = [S(abi.SENSOR_ACCELEROMETER, tree=0, link=1)]
    with pytest.raises(abi.NewtError) as e:
        Model(bad)
    assert e.value.status == abi.ERR_UNSUPPORTED
