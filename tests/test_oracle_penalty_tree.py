"""Penalty contacts on a tree (tree_wrenches_from_pairs and
apply_tree_contact_wrench, newt/src/world.rs:2570-2741) in the oracle,
pinned by closed forms: no golden of newt's holds a tree under penalty
contacts.  A sphere on a vertical slide joint over a plane: one step's qacc
is (k pen - c v + m g) / m with k = m / tc^2 and c = 2 dampratio m / tc
(solref_to_kc, geom.rs:137-145); at rest the spring carries the weight,
pen = m g / k = g tc^2.
"""
import numpy as np
import pytest

from tooling_amd import abi, scenes


def _slide_sphere(integrator=abi.INTEGRATOR_EULER, dt=0.001, mass=1.7, radius=0.1, solref=(0.02, 1.0)):
    links = [scenes.link(-1, abi.JOINT_FIXED, mass=1.0, inertia=(1, 0, 0, 0, 1, 0, 0, 0, 1)),
             scenes.link(0, abi.JOINT_SLIDE, axis=(0.0, 0.0, 1.0), mass=mass,
                         inertia=(0.01, 0, 0, 0, 0.01, 0, 0, 0, 0.01))]
    geoms = [scenes.geom(abi.SHAPE_PLANE, friction=0.5, solref=solref),
             scenes.geom(abi.SHAPE_SPHERE, tree=0, link=1, size=(radius, 0.0, 0.0), friction=0.5, solref=solref)]
    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=integrator, solver=abi.SOLVER_PENALTY, iterations=1)
    return scenes.Scene("slide_sphere", opts, links, geoms, 0, None)


@pytest.mark.parametrize("q,qd", [(0.08, 0.0), (0.08, -0.5), (0.095, 0.3), (0.2, -1.0)])
def test_slide_sphere_one_step_closed_form(oracle_lib, q, qd):
    m, tc, dr, r = 1.7, 0.02, 1.0, 0.1
    ow = oracle_lib.OracleWorlds(_slide_sphere(), 1, 0)
    ow.set_state(np.array([[q]]), np.array([[qd]]))
    ow.step(1)
    a = ow.qacc()[0, 0]
    ow.close()
    k, c = m / (tc * tc), 2.0 * dr * m / tc
    pen = r - q
    fn = max(k * pen - c * qd, 0.0) if pen > 0.0 else 0.0
    assert a == pytest.approx((fn - m * 9.81) / m, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_EULER, abi.INTEGRATOR_RK4])
def test_slide_sphere_rests_on_its_spring(oracle_lib, integrator):
    ow = oracle_lib.OracleWorlds(_slide_sphere(integrator), 1, 0)
    ow.set_state(np.array([[0.3]]), np.zeros((1, 1)))
    ow.step(4000)
    q, v = ow.state()
    ow.close()
    assert abs(v[0, 0]) < 1e-6
    assert (0.1 - q[0, 0]) == pytest.approx(9.81 * 0.02 ** 2, rel=1e-6)


def _free_sphere(mass, radius, inertia, mu, solref=(0.02, 1.0), dt=0.001):
    """A free-root tree of one sphere link over a plane (the shape of
    free_sphere_tree, newt/tests/tree_contacts_solver.rs:17)."""
    links = [scenes.link(-1, abi.JOINT_FREE, mass=mass, inertia=(inertia, 0, 0, 0, inertia, 0, 0, 0, inertia))]
    geoms = [scenes.geom(abi.SHAPE_PLANE, friction=mu, solref=solref),
             scenes.geom(abi.SHAPE_SPHERE, tree=0, link=0, size=(radius, 0.0, 0.0), friction=mu, solref=solref)]
    opts = dict(dt=dt, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_PENALTY,
                iterations=1)
    return scenes.Scene("free_sphere", opts, links, geoms, 0, None)


@pytest.mark.parametrize("vx,wx", [(0.3, 2.0), (0.01, 0.1), (-0.02, 6.0)])
def test_free_sphere_friction_and_torque_closed_form(oracle_lib, vx, wx):
    """A sliding, spinning sphere pressed 5 mm into the plane: the contact
    force is the spring k·pen upward plus, per tangent axis, the damper
    −c·v_slip clamped to ±μ·f_n (the pyramidal clamp of
    apply_tree_contact_wrench, world.rs:2633-2741; for a +z normal newt's
    tangent basis is (−y, +x), world.rs:3104-3111, so the clamp acts per world
    axis).  The slip velocity is the contact point's, v + ω × r, and the
    torque about the COM is r × F, with r from the COM to the contact point
    (midway through the overlap, contact.rs:269-301).  With ω ∥ v and an
    isotropic inertia, Newton-Euler in body coordinates gives
    q̈_ang = (r × F)/I and q̈_lin = F/m + g."""
    m, r, inertia, mu, tc, dr = 2.0, 0.1, 0.008, 0.5, 0.02, 1.0
    h = 0.095
    ow = oracle_lib.OracleWorlds(_free_sphere(m, r, inertia, mu), 1, 0)
    q = np.array([[0.0, 0.0, h, 0.0, 0.0, 0.0, 1.0]])
    v = np.array([[wx, 0.0, 0.0, vx, 0.0, 0.0]])
    ow.set_state(q, v)
    ow.step(1)
    a = ow.qacc()[0]
    ow.close()
    pen = r - h
    arm = np.array([0.0, 0.0, (h - (r - pen * 0.5)) - h])
    slip = np.array([vx, 0.0, 0.0]) + np.cross([wx, 0.0, 0.0], arm)
    k, c = m / tc ** 2, 2.0 * dr * m / tc
    fn = k * pen
    ft = np.clip(-c * slip[:2], -mu * fn, mu * fn)
    force = np.array([ft[0], ft[1], fn])
    np.testing.assert_allclose(a[0:3], np.cross(arm, force) / inertia, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(a[3:6], force / m + np.array([0.0, 0.0, -9.81]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("qd1,qd2", [(0.2, -0.1), (0.0, 0.0), (-0.3, 0.5)])
def test_self_contact_equal_and_opposite_closed_form(oracle_lib, qd1, qd2):
    """Two sphere links on sibling slide joints (along x) of one tree,
    overlapping by 5 cm, self collision on: the pair's spring-damper uses the
    reduced mass m1 m2/(m1 + m2) (apply_tree_contact_wrench,
    world.rs:2660-2690), and the two links receive equal and opposite
    forces along the line of centres: m1 q̈1 = −m2 q̈2 = −f_n with
    f_n = max(k pen + c (q̇1 − q̇2), 0)."""
    m1, m2, r, tc, dr = 1.0, 3.0, 0.1, 0.02, 1.0
    inert = (0.01, 0, 0, 0, 0.01, 0, 0, 0, 0.01)
    links = [scenes.link(-1, abi.JOINT_FIXED, mass=1.0, inertia=inert),
             scenes.link(0, abi.JOINT_SLIDE, axis=(1.0, 0.0, 0.0), mass=m1, inertia=inert),
             scenes.link(0, abi.JOINT_SLIDE, axis=(1.0, 0.0, 0.0), mass=m2, inertia=inert)]
    geoms = [scenes.geom(abi.SHAPE_SPHERE, tree=0, link=1, size=(r, 0.0, 0.0)),
             scenes.geom(abi.SHAPE_SPHERE, tree=0, link=2, size=(r, 0.0, 0.0))]
    opts = dict(dt=0.001, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_EULER, solver=abi.SOLVER_PENALTY,
                iterations=1)
    sc = scenes.Scene("self_pair", opts, links, geoms, 1, None)
    ow = oracle_lib.OracleWorlds(sc, 1, 0)
    ow.set_state(np.array([[0.0, 0.15]]), np.array([[qd1, qd2]]))
    ow.step(1)
    a = ow.qacc()[0]
    ow.close()
    me = m1 * m2 / (m1 + m2)
    k, c = me / tc ** 2, 2.0 * dr * me / tc
    fn = max(k * (2 * r - 0.15) + c * (qd1 - qd2), 0.0)
    assert m1 * a[0] == pytest.approx(-fn, rel=1e-12, abs=1e-12)
    assert m2 * a[1] == pytest.approx(fn, rel=1e-12, abs=1e-12)
