"""Pins of the oracle's convex-mesh contacts (oracle/newt_ccd.hpp: GJK, EPA,
the mesh-mesh manifold, mesh_plane, sphere_mesh), on the CPU.

- newt's own fixture test, analytic_convex_route_probes_are_fixture_backed
  (newt/tests/contacts_geoms_v1.rs:1575-1693), restated: MuJoCo 3.11's
  mj_forward contacts for sphere-mesh, plane-mesh, mesh-mesh and a mesh-mesh
  face manifold at 20 poses (tests/golden/contact_route_probes.json, a copy
  of newt/tests/references/contact_route_probes.json) must be met in count,
  and in position, normal and penetration within the reviewed bounds of
  contact_route_probe_bounds.json, by the float instantiation (newt's
  arithmetic, which the probes were reviewed against).  The box-mesh probe
  is skipped, as newt skips it (box-mesh is not a supported pair).  The
  double instantiation is not held to the probes: where a mesh face lies
  flat on the plane, mesh_plane's second vertex is picked by a tangent
  coordinate that differs between tied vertices only by rounding, so float
  and double may pick different (equally deep) vertices.
- closed forms: a unit cube mesh resting face-down on another (the manifold
  is the four corners of the overlap square at penetration = overlap depth,
  normal +Z), a cube pushed into a plane (two support-vertex contacts at the
  penetration depth, midpoints half way), a sphere against a cube face and
  against a cube edge (penetration r - d, normal along the face normal or
  the edge's outward bisector), separated meshes give nothing, and both
  instantiations agree to float rounding.
- argument order: swapping the two meshes flips the normal and keeps the
  points (newt's mesh_mesh_witness_is_invariant_to_vertex_order_and_argument
  _order, contacts_geoms_v1.rs:453-521, and
  coincident_mesh_mesh_swapped_order_flips_normal, 562-588).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle
from tooling_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module", autouse=True)
def _lib():
    pyoracle.lib()


def _geom(shape, size=(0.0, 0.0, 0.0), margin=0.0):
    g = abi.GeomDesc()
    g.shape = shape
    g.size = abi.D3(*size)
    g.friction = 0.5
    g.margin = margin
    g.mesh = 0
    return g


def _pose(d):
    p, q = d["position"], d["orientation_wxyz"]
    return (p[0], p[1], p[2], q[1], q[2], q[3], q[0])


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


PROBE_GEOMS = {
    "sphere-mesh": (lambda: _geom(abi.SHAPE_SPHERE, (0.2, 0.0, 0.0)), lambda: _geom(abi.SHAPE_MESH)),
    "plane-mesh": (lambda: _geom(abi.SHAPE_PLANE), lambda: _geom(abi.SHAPE_MESH)),
    "mesh-mesh": (lambda: _geom(abi.SHAPE_MESH), lambda: _geom(abi.SHAPE_MESH)),
    "mesh-mesh-manifold": (lambda: _geom(abi.SHAPE_MESH), lambda: _geom(abi.SHAPE_MESH)),
}


def _probe_cases():
    doc = _load("contact_route_probes.json")
    out = []
    for probe in doc["probes"]:
        if probe["pair"] not in PROBE_GEOMS:
            continue
        for pose in probe["poses"]:
            out.append(pytest.param(probe["id"], probe["pair"], pose["id"], id=f"{probe['id']}-{pose['id']}"))
    return out


@pytest.mark.parametrize("probe_id,pair,pose_id", _probe_cases())
def test_mujoco_route_probes(probe_id, pair, pose_id, precision=1):
    doc = _load("contact_route_probes.json")
    bounds = {b["id"]: b["bounds"] for b in _load("contact_route_probe_bounds.json")["probes"]}[probe_id]
    probe = next(p for p in doc["probes"] if p["id"] == probe_id)
    pose = next(p for p in probe["poses"] if p["id"] == pose_id)
    mesh = None
    if probe.get("mesh"):
        mesh = (np.array(probe["mesh"]["vertices"], dtype=np.float64), np.array(probe["mesh"]["faces"], dtype=np.int32))
    ga, gb = PROBE_GEOMS[pair]
    got = pyoracle.narrow_phase(ga(), _pose(pose["pose_a"]), gb(), _pose(pose["pose_b"]), mesh, precision)
    expected = pose["contacts"]
    count = 0 if not expected else len(expected) + int(bounds["count_delta"])
    assert len(got) == count, (probe_id, pose_id, got)
    for k, e in enumerate(expected):
        pos = np.array(e["position"])
        nrm = -np.array(e["frame_normal"])
        assert np.linalg.norm(got[k, :3] - pos) <= bounds["position"], (probe_id, pose_id, k, got[k], e)
        assert np.linalg.norm(got[k, 3:6] - nrm) <= bounds["normal"], (probe_id, pose_id, k, got[k], e)
        assert abs(got[k, 6] - e["penetration"]) <= bounds["penetration"], (probe_id, pose_id, k, got[k], e)


def cube(h=0.5):
    """Unit cube (half extent h) with outward-wound faces."""
    v = np.array([[-h, -h, -h], [h, -h, -h], [-h, h, -h], [h, h, -h],
                  [-h, -h, h], [h, -h, h], [-h, h, h], [h, h, h]])
    f = np.array([[0, 3, 1], [0, 2, 3], [4, 5, 7], [4, 7, 6], [0, 1, 5], [0, 5, 4],
                  [2, 6, 7], [2, 7, 3], [0, 4, 6], [0, 6, 2], [1, 3, 7], [1, 7, 5]], dtype=np.int32)
    return v, f


ID = (0.0, 0.0, 0.0, 1.0)


@pytest.mark.parametrize("precision", [1, 0])
def test_cube_on_cube_face_manifold(precision):
    depth, dx, dy = 0.03, 0.2, 0.1
    got = pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), (dx, dy, 1.0 - depth) + ID, _geom(abi.SHAPE_MESH),
                                (0.0, 0.0, 0.0) + ID, cube(), precision)
    assert len(got) == 4
    tol = 2e-6 if precision == 1 else 1e-12
    np.testing.assert_allclose(got[:, 3:6], np.tile([0.0, 0.0, 1.0], (4, 1)), atol=tol)
    np.testing.assert_allclose(got[:, 6], depth, atol=tol)
    # corners of the overlap square [dx - 0.5, 0.5] x [dy - 0.5, 0.5], at
    # the midplane of the overlap
    xy = {(round(x, 5), round(y, 5)) for x, y in got[:, :2]}
    assert xy == {(round(x, 5), round(y, 5)) for x in (dx - 0.5, 0.5) for y in (dy - 0.5, 0.5)}
    np.testing.assert_allclose(got[:, 2], 0.5 - depth / 2, atol=tol)


@pytest.mark.parametrize("precision", [1, 0])
def test_cube_into_plane(precision):
    depth = 0.04
    tilt = 0.3
    q = (0.0, math.sin(tilt / 2), 0.0, math.cos(tilt / 2))
    # the lowest edge of the tilted cube: two tied support vertices
    v, f = cube()
    R = np.array([[math.cos(tilt), 0.0, math.sin(tilt)], [0.0, 1.0, 0.0], [-math.sin(tilt), 0.0, math.cos(tilt)]])
    low = (v @ R.T)[:, 2].min()
    got = pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), (0.0, 0.0, -low - depth) + q, _geom(abi.SHAPE_PLANE),
                                (0.0, 0.0, 0.0) + ID, (v, f), precision)
    assert len(got) == 2
    tol = 1e-5 if precision == 1 else 1e-12
    np.testing.assert_allclose(got[:, 6], depth, atol=tol)
    np.testing.assert_allclose(got[:, 3:6], np.tile([0.0, 0.0, 1.0], (2, 1)), atol=tol)
    np.testing.assert_allclose(got[:, 2], -depth / 2, atol=tol)
    assert sorted(np.round(got[:, 1], 5)) == [-0.5, 0.5]  # primary: largest y; secondary: smallest


@pytest.mark.parametrize("precision", [1, 0])
def test_sphere_against_cube_face_and_edge(precision):
    r = 0.2
    tol = 1e-6 if precision == 1 else 1e-12
    got = pyoracle.narrow_phase(_geom(abi.SHAPE_SPHERE, (r, 0, 0)), (0.1, -0.2, 0.5 + r - 0.05) + ID,
                                _geom(abi.SHAPE_MESH), (0.0, 0.0, 0.0) + ID, cube(), precision)
    assert len(got) == 1
    np.testing.assert_allclose(got[0, 3:6], [0.0, 0.0, 1.0], atol=tol)
    np.testing.assert_allclose(got[0, 6], 0.05, atol=tol)
    np.testing.assert_allclose(got[0, :3], [0.1, -0.2, 0.5 - 0.025], atol=tol)
    d = 0.15
    c = 0.5 + d / math.sqrt(2.0)
    got = pyoracle.narrow_phase(_geom(abi.SHAPE_SPHERE, (r, 0, 0)), (c, 0.0, c) + ID, _geom(abi.SHAPE_MESH),
                                (0.0, 0.0, 0.0) + ID, cube(), precision)
    assert len(got) == 1
    np.testing.assert_allclose(got[0, 3:6], [1 / math.sqrt(2), 0.0, 1 / math.sqrt(2)], atol=tol)
    np.testing.assert_allclose(got[0, 6], r - d, atol=tol)


@pytest.mark.parametrize("precision", [1, 0])
def test_separated_meshes_and_argument_order(precision):
    assert len(pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), (1.2, 0.3, 0.1) + ID, _geom(abi.SHAPE_MESH),
                                     (0.0, 0.0, 0.0) + ID, cube(), precision)) == 0
    q = (0.1, 0.2, -0.15)
    n = math.sqrt(1.0 - sum(x * x for x in q))
    pa = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    pb = (0.7, 0.35, 0.4) + q + (n,)
    ab = pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), pa, _geom(abi.SHAPE_MESH), pb, cube(), precision)
    ba = pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), pb, _geom(abi.SHAPE_MESH), pa, cube(), precision)
    assert len(ab) >= 1 and len(ab) == len(ba)
    tol = 1e-5 if precision == 1 else 1e-12
    np.testing.assert_allclose(ab[:, 3:6], -ba[:, 3:6], atol=tol)
    np.testing.assert_allclose(ab[:, 6], ba[:, 6], atol=tol)
    key = lambda a: sorted(map(tuple, np.round(a[:, :3], 4)))
    assert key(ab) == key(ba)


def test_float_and_double_agree():
    r = np.random.default_rng(7)
    worst = 0.0
    for _ in range(40):
        q = r.normal(size=4)
        q /= np.linalg.norm(q)
        pb = tuple(r.uniform(-0.9, 0.9, size=3)) + tuple(q)
        a = pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), (0.0,) * 6 + (1.0,), _geom(abi.SHAPE_MESH), pb, cube(), 1)
        b = pyoracle.narrow_phase(_geom(abi.SHAPE_MESH), (0.0,) * 6 + (1.0,), _geom(abi.SHAPE_MESH), pb, cube(), 0)
        if len(a) == len(b) and len(a):
            worst = max(worst, float(np.abs(a - b).max()))
    assert worst < 1e-4


TETRA = (np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
         np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.int32))


def tetra_tumble_scene(rotated):
    """build_dynamic_anchor_world (newt/tests/contacts_geoms_v1.rs:2194-2229):
    a static unit tetrahedron and a free one of mass 1 (principal inertia
    0.166667) overlapping it, zero gravity, dt 5 ms, World::new's RK4 and
    penalty contacts."""
    from tooling_amd import scenes
    if rotated:
        pos, q, w = (0.25, 0.1, 0.25), scenes.quat_axis_angle((0.0, 0.0, 1.0), math.radians(0.4)), (-0.6, 0.9, 0.5)
    else:
        pos, q, w = (0.2, 0.2, 0.2), (0.0, 0.0, 0.0, 1.0), (1.0, 0.7, -0.4)
    i = 0.166667
    b = scenes.body(1.0, (i, 0.0, 0.0, 0.0, i, 0.0, 0.0, 0.0, i), pos, q, angular_velocity_body=w)
    geoms = [scenes.geom(abi.SHAPE_MESH, mesh=0, friction=0.5), scenes.geom(abi.SHAPE_MESH, body=0, mesh=0, friction=0.5)]
    qpos = np.array([list(pos) + list(q)])
    qvel = np.array([[0.0, 0.0, 0.0] + list(w)])
    opts = dict(dt=0.005, gravity=(0.0, 0.0, 0.0), integrator=abi.INTEGRATOR_RK4, solver=abi.SOLVER_PENALTY,
                iterations=20)
    return scenes.Scene("tetra_tumble" + ("_rotated" if rotated else ""), opts, [], geoms, 0,
                        lambda n, seed: (np.repeat(qpos, n, 0), np.repeat(qvel, n, 0)), "", bodies=[b],
                        meshes=[TETRA])


@pytest.mark.parametrize("case_id", ["mesh-mesh-tumble", "mesh-mesh-rotated-drop"])
def test_mujoco_dynamic_mesh_anchors(case_id):
    """newt's dynamic_enabled_convex_anchors_are_fixture_backed
    (contacts_geoms_v1.rs:2262-2420): MuJoCo's mj_step trajectory of the
    tumbling tetrahedron over 100 steps (tests/golden/contact_dynamic_anchors
    .json) within the reviewed bounds of contact_dynamic_anchor_bounds.json
    (position and orientation error per window, contact count exact), on the
    float instantiation stepping World::step."""
    doc = _load("contact_dynamic_anchors.json")
    bounds = next(c for c in _load("contact_dynamic_anchor_bounds.json")["cases"] if c["id"] == case_id)
    case = next(c for c in doc["cases"] if c["id"] == case_id)
    sc = tetra_tumble_scene(case_id == "mesh-mesh-rotated-drop")
    ow = pyoracle.OracleWorlds(sc, 1, 1)
    ow.set_state(*sc.initial_state(1, 0))
    done = 0
    worst = {"early": [0.0, 0.0, 0], "full": [0.0, 0.0, 0]}
    for sample in case["samples"]:
        while done < sample["step"]:
            ow.step(1, threads=1)
            done += 1
        q = ow.state()[0][0]
        ep = np.array(sample["position"])
        w, x, y, z = sample["orientation_wxyz"]
        perr = float(np.linalg.norm(q[:3] - ep))
        oerr = float(np.linalg.norm(q[3:7] - np.array([x, y, z, w])))
        cerr = abs(ow.detect(0) - sample["contacts"])
        for win in (["early", "full"] if done <= 20 else ["full"]):
            b = bounds[win]
            worst[win] = [max(worst[win][0], perr), max(worst[win][1], oerr), max(worst[win][2], cerr)]
            assert perr <= b["position"], (case_id, done, perr)
            assert oerr <= b["orientation"], (case_id, done, oerr)
            assert cerr <= b["contact_count"], (case_id, done, ow.detect(0), sample["contacts"])
    print(case_id, worst)
    ow.close()
