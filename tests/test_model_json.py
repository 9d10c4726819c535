"""The JSON scene loader (newt's model format, docs/model-format.md).

Strictness and error paths follow newt/src/model.rs; the fixtures in
tests/golden/models are newt's own models/pendulum.json, arm.json and
stack.json (copied as data) and quadruped18.json, written by
tests/golden/make_model_json.py from the programmatic quadruped.
"""
import json
import os

import numpy as np
import pytest

from tooling_amd import abi, model_json, scenes

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")


def test_newt_pendulum_model_loads():
    sc = model_json.load(os.path.join(MODELS, "pendulum.json"))
    assert sc.options["integrator"] == abi.INTEGRATOR_RK4          # format default
    assert sc.options["solver"] == abi.SOLVER_PENALTY and sc.options["iterations"] == 20
    assert sc.options["dt"] == float(np.float32(0.005))
    assert [l["joint"] for l in sc.links] == [abi.JOINT_FIXED, abi.JOINT_HINGE, abi.JOINT_HINGE]
    assert [l["parent"] for l in sc.links] == [-1, 0, 1]
    assert sc.links[0]["parent_pos"] == (0.0, 0.0, 1.5)
    assert sc.links[2]["child_pos"] == pytest.approx((0.0, 0.0, 0.3))
    assert sc.links[1]["mass"] == pytest.approx(1.3) and sc.links[1]["inertia"][0] == float(np.float32(0.08775))
    assert sc.extra["sites"][0]["name"] == "tip"


@pytest.mark.parametrize("name,path", [("arm", "$.actuators")])
def test_unbuilt_features_are_rejected_with_a_path(name, path):
    with pytest.raises(model_json.ModelError) as e:
        model_json.load(os.path.join(MODELS, f"{name}.json"))
    assert e.value.path == path


def _pendulum_doc():
    with open(os.path.join(MODELS, "pendulum.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("mutate,path", [
    (lambda d: d["trees"][0]["links"][1]["joint"].update(dampign=0.1), "$.trees[0].links[1].joint.dampign"),
    (lambda d: d["trees"][0]["links"][1]["joint"].update(axis=[0, 0, 0]), "$.trees[0].links[1].joint.axis"),
    (lambda d: d["trees"][0]["links"][1]["joint"].update(kind="ball"), "$.trees[0].links[1].joint.kind"),
    (lambda d: d["trees"][0]["links"][2].update(parent="nowhere"), "$.trees[0].links[2].parent"),
    (lambda d: d["trees"][0]["links"][1]["joint"].update(range=[1.0, -1.0]), "$.trees[0].links[1].joint.range"),
    (lambda d: d["trees"][0]["links"][1].update(mass=-1.0), "$.trees[0].links[1].mass"),
    (lambda d: d.update(integrator="Verlet"), "$.integrator"),
    (lambda d: d.update(geoms=[{"shape": {"kind": "sphere", "radius": 0.1}}]), "$.geoms[0].attach"),
    (lambda d: d.update(geoms=[{"shape": {"kind": "cylinder", "radius": 0.1, "half_height": 0.2},
                                "attach": {"kind": "link", "tree": "pendulum", "link": "upper"}}]),
     "$.geoms[0].shape.kind"),
])
def test_strict_errors_name_the_offending_value(mutate, path):
    d = _pendulum_doc()
    mutate(d)
    with pytest.raises(model_json.ModelError) as e:
        model_json.load(json.dumps(d))
    assert e.value.path == path


def test_duplicate_keys_are_rejected():
    with pytest.raises(model_json.ModelError):
        model_json.load('{"timestep": 0.01, "timestep": 0.02, "trees": []}')


def test_limits_and_solver_fields():
    d = _pendulum_doc()
    d["solver"] = {"mode": "pgs", "iterations": 30}
    d["trees"][0]["links"][1]["joint"].update(range=[-0.5, 0.5], damping=0.25, armature=0.01,
                                              limit={"stiffness": 2000, "damping": 40,
                                                     "solref": {"timeconst": 0.05, "dampratio": 0.9}})
    sc = model_json.load(json.dumps(d))
    l1 = sc.links[1]
    assert l1["range"] == (-0.5, 0.5) and l1["damping"] == float(np.float32(0.25))
    assert l1["limit_solref"] == pytest.approx((0.05, 0.9)) and l1["limit_solimp"] is None
    assert l1["limit_stiffness"] == 2000.0 and l1["limit_damping"] == 40.0
    assert sc.options["solver"] == abi.SOLVER_PGS and sc.options["iterations"] == 30


def test_quadruped_json_matches_the_programmatic_scene():
    """Every link and geom field the step reads, within float32 rounding
    (newt parses JSON numbers as f32)."""
    js = model_json.load(os.path.join(MODELS, "quadruped18.json"))
    pr = scenes.quadruped()
    assert len(js.links) == len(pr.links) and len(js.geoms) == len(pr.geoms)
    for a, b in zip(js.links, pr.links):
        assert a["parent"] == b["parent"] and a["joint"] == b["joint"]
        for k in ("axis", "parent_pos", "parent_quat", "child_pos", "inertia"):
            assert np.allclose(a[k], b[k], rtol=1e-6, atol=1e-7), k
        assert a["mass"] == pytest.approx(b["mass"], rel=1e-7)
    for a, b in zip(js.geoms, pr.geoms):
        assert a["shape"] == b["shape"] and a["link"] == b["link"] and a["condim"] == b["condim"]
        for k in ("size", "pos", "quat", "solref", "solimp"):
            assert np.allclose(a[k], b[k], rtol=1e-6, atol=1e-7), k
        assert a["friction"] == pytest.approx(b["friction"], rel=1e-7)
    for k in ("integrator", "solver", "iterations"):
        assert js.options[k] == pr.options[k]


def test_newt_stack_model_loads_free_bodies(oracle_lib):
    """newt's stack.json (newt/tests/model_load.rs:202-221): three solid
    boxes of half extent 0.35 with masses 1.2, 0.9, 1.5, the top one spinning
    at 0.3 rad/s about body y, box geoms on the bodies, a static plane, all
    in float32 as newt parses them.  Stepped by the oracle's float
    instantiation under PGS, the loaded scene and the same scene built from
    the reference's programmatic description give the same bytes."""
    sc = model_json.load(os.path.join(MODELS, "stack.json"))
    assert len(sc.links) == 0 and len(sc.bodies) == 3 and sc.extra["body_names"] == ["bottom", "middle", "top"]
    f32 = lambda x: float(np.float32(x))
    half = (0.35, 0.35, 0.35)
    expect = [(1.2, (0.0, 0.0, 0.5), (0.0, 0.0, 0.0)), (0.9, (0.02, 0.0, 1.7), (0.0, 0.0, 0.0)),
              (1.5, (0.0, 0.0, 2.9), (0.0, 0.3, 0.0))]
    prog = []
    for b, (m, p, w) in zip(sc.bodies, expect):
        assert b["mass"] == f32(m)
        assert b["position"] == tuple(f32(x) for x in p)
        assert b["angular_velocity_body"] == tuple(f32(x) for x in w)
        m32, h32 = np.float32(m), np.float32(0.35)
        t = np.float32(m32 / np.float32(3.0))
        i = float(np.float32(t * np.float32(h32 * h32 + h32 * h32)))
        assert b["inertia"] == (i, 0, 0, 0, i, 0, 0, 0, i)
        prog.append(scenes.body(f32(m), (i, 0, 0, 0, i, 0, 0, 0, i), tuple(f32(x) for x in p),
                                angular_velocity_body=tuple(f32(x) for x in w)))
    assert [g["body"] for g in sc.geoms] == [-1, 0, 1, 2]
    assert sc.options["solver"] == abi.SOLVER_PENALTY
    geoms = [scenes.geom(abi.SHAPE_PLANE, friction=f32(0.6))]
    geoms += [scenes.geom(abi.SHAPE_BOX, body=k, size=tuple(f32(x) for x in half), friction=f32(0.6))
              for k in range(3)]
    pgs = dict(sc.options, solver=abi.SOLVER_PGS, iterations=30)
    loaded = scenes.Scene("loaded", pgs, [], sc.geoms, 0, None, bodies=sc.bodies)
    built = scenes.Scene("built", pgs, [], geoms, 0, None, bodies=prog)
    out = []
    for s2 in (loaded, built):
        ow = oracle_lib.OracleWorlds(s2, 1, 1)
        ow.step(500)
        out.append(np.concatenate(ow.state(), axis=1).astype(np.float32).tobytes())
        ow.close()
    assert out[0] == out[1]


def test_body_attachment_errors():
    d = json.load(open(os.path.join(MODELS, "stack.json")))
    d["geoms"][1]["attach"]["body"] = "nowhere"
    with pytest.raises(model_json.ModelError) as e:
        model_json.load(json.dumps(d))
    assert e.value.path == "$.geoms[1].attach.body"
    d = json.load(open(os.path.join(MODELS, "stack.json")))
    d["bodies"][0]["rolling_friction"] = 0.1
    with pytest.raises(model_json.ModelError) as e:
        model_json.load(json.dumps(d))
    assert e.value.path == "$.bodies[0].rolling_friction"
