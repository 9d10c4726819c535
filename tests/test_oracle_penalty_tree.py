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
