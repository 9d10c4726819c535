"""Closed-form pins of the oracle's sphere/capsule contacts
(sphere_sphere, sphere_capsule, capsule_capsule and sphere_vs_sphere_at,
newt/src/contact.rs:407-499, 1249-1272, 2313-2385).

One free body carries geom 0; geom 1 is static.  The contact of the first
step (detected at the initial state) is compared with the geometry worked
by hand: the normal points from geom 1 into geom 0, the point lies on geom
1's surface, the penetration is r0 + r1 - distance.  The segment cases are
the reference's own unit tests (contact.rs:3546-3585): clamping past the
ends, perpendicular segments meeting at their midpoints, parallel segments
separated by their common perpendicular.
"""
import math

import numpy as np
import pytest

from tooling_amd import abi, scenes

S2 = math.sqrt(0.5)
Z_TO_X = (0.0, S2, 0.0, S2)    # rotates local +Z onto +X
Z_TO_Y = (-S2, 0.0, 0.0, S2)   # rotates local +Z onto +Y


def _contact(oracle_lib, g0, g1, body_pos):
    links = [scenes.link(-1, abi.JOINT_FREE, mass=1.0, inertia=(1, 0, 0, 0, 1, 0, 0, 0, 1))]
    geoms = [dict(g0, tree=0, link=0), g1]
    opts = dict(dt=0.001, gravity=(0.0, 0.0, 0.0), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_PGS,
                iterations=1)
    sc = scenes.Scene("pair", opts, links, geoms, 0, None)
    ow = oracle_lib.OracleWorlds(sc, 1, 0)
    ow.set_state(np.array([[*body_pos, 0.0, 0.0, 0.0, 1.0]]), np.zeros((1, 6)))
    ow.step(1)
    ids, geo = ow.contacts(0)
    ow.close()
    return ids, geo


def _check(ids, geo, point, normal, pen):
    assert len(ids) == 1 and tuple(ids[0]) == (0, 1, 0)
    assert np.allclose(geo[0, 0:3], point, atol=1e-12)
    assert np.allclose(geo[0, 3:6], normal, atol=1e-12)
    assert geo[0, 6] == pytest.approx(pen, abs=1e-12)


def test_sphere_sphere(oracle_lib):
    c0, c1, r0, r1 = np.array([0.3, -0.2, 1.0]), np.array([0.0, 0.1, 0.6]), 0.25, 0.4
    ids, geo = _contact(oracle_lib, scenes.geom(abi.SHAPE_SPHERE, size=(r0, 0, 0)),
                        scenes.geom(abi.SHAPE_SPHERE, size=(r1, 0, 0), pos=tuple(c1)), c0)
    d = np.linalg.norm(c0 - c1)
    n = (c0 - c1) / d
    _check(ids, geo, c1 + n * r1, n, r0 + r1 - d)


def test_sphere_sphere_coincident_centers_take_plus_z(oracle_lib):
    ids, geo = _contact(oracle_lib, scenes.geom(abi.SHAPE_SPHERE, size=(0.2, 0, 0)),
                        scenes.geom(abi.SHAPE_SPHERE, size=(0.3, 0, 0), pos=(1.0, 2.0, 3.0)), (1.0, 2.0, 3.0))
    _check(ids, geo, (1.0, 2.0, 3.3), (0.0, 0.0, 1.0), 0.5)


def test_sphere_sphere_apart(oracle_lib):
    ids, _ = _contact(oracle_lib, scenes.geom(abi.SHAPE_SPHERE, size=(0.2, 0, 0)),
                      scenes.geom(abi.SHAPE_SPHERE, size=(0.3, 0, 0)), (0.0, 0.0, 0.51))
    assert len(ids) == 0


@pytest.mark.parametrize("x", [0.35, 2.5, -1.7])
def test_sphere_capsule_interior_and_clamped(oracle_lib, x):
    """Static capsule along +X (half height 1, radius 0.2) at the origin;
    the sphere's closest axis point is its projection, clamped to the ends."""
    c0, r0 = np.array([x, 0.1, 0.25]), 0.15
    ids, geo = _contact(oracle_lib, scenes.geom(abi.SHAPE_SPHERE, size=(r0, 0, 0)),
                        scenes.geom(abi.SHAPE_CAPSULE, size=(0.2, 1.0, 0), quat=Z_TO_X), c0)
    axis_pt = np.array([min(max(x, -1.0), 1.0), 0.0, 0.0])
    d = np.linalg.norm(c0 - axis_pt)
    n = (c0 - axis_pt) / d
    if r0 + 0.2 - d <= 0.0:
        assert len(ids) == 0
        return
    _check(ids, geo, axis_pt + n * 0.2, n, r0 + 0.2 - d)


def test_capsule_sphere_is_relabelled(oracle_lib):
    """(capsule, sphere) runs as sphere-capsule, then the labels swap and the
    normal flips: it still points from geom 1 (the sphere) into geom 0."""
    cs, rs = np.array([0.4, 0.0, 0.3]), 0.2
    ids, geo = _contact(oracle_lib, scenes.geom(abi.SHAPE_CAPSULE, size=(0.15, 0.5, 0), quat=Z_TO_X),
                        scenes.geom(abi.SHAPE_SPHERE, size=(rs, 0, 0), pos=tuple(cs)), (0.0, 0.0, 0.0))
    axis_pt = np.array([0.4, 0.0, 0.0])
    d = np.linalg.norm(cs - axis_pt)
    n_sc = (cs - axis_pt) / d               # sphere_vs_sphere_at(sphere, capsule point)
    _check(ids, geo, axis_pt + n_sc * 0.15, -n_sc, rs + 0.15 - d)


def test_capsule_capsule_perpendicular(oracle_lib):
    """contact.rs:3559-3570: segments along X at z=0 and along Y at z=1
    meet at their midpoints; radii 0.6 and 0.5 overlap by 0.1."""
    ids, geo = _contact(oracle_lib, scenes.geom(abi.SHAPE_CAPSULE, size=(0.6, 1.0, 0), quat=Z_TO_X),
                        scenes.geom(abi.SHAPE_CAPSULE, size=(0.5, 1.0, 0), pos=(0.0, 0.0, 1.0), quat=Z_TO_Y),
                        (0.0, 0.0, 0.0))
    _check(ids, geo, (0.0, 0.0, 0.5), (0.0, 0.0, -1.0), 0.1)


def test_capsule_capsule_parallel(oracle_lib):
    """contact.rs:3572-3585: parallel segments [0,1]x{0} and [0.3,1.3]x{0.7};
    the separation is perpendicular to both and 0.7 long."""
    ids, geo = _contact(oracle_lib, scenes.geom(abi.SHAPE_CAPSULE, size=(0.4, 0.5, 0), quat=Z_TO_X),
                        scenes.geom(abi.SHAPE_CAPSULE, size=(0.4, 0.5, 0), pos=(0.8, 0.7, 0.0), quat=Z_TO_X),
                        (0.5, 0.0, 0.0))
    assert len(ids) == 1
    n = geo[0, 3:6]
    assert np.allclose(n, (0.0, -1.0, 0.0), atol=1e-12)
    assert geo[0, 6] == pytest.approx(0.8 - 0.7, abs=1e-12)
    # the point lies on geom 1's surface, between the overlapping ends
    p = geo[0, 0:3]
    assert p[1] == pytest.approx(0.3, abs=1e-12) and 0.3 - 1e-12 <= p[0] <= 1.0 + 1e-12
