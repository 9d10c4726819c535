"""Trees and free bodies in one world, on the oracle.  A contact between a
tree link and a free body enters the tree solve with a body component:
world_jacobian_for_side, apply_world_impulse, world_velocity_from_delta,
world_current_velocity, world_free_velocity, tree_contact_diag_approx and
the body_wrenches recovery (newt/src/solver.rs:2805-3366, 3537-3836).  The
body step adds the tree solve's body wrenches to those of its own solve
(World::compute_solver_wrenches, newt/src/world.rs:2097-2149).

Pinned by:
- newt's anchors in newt/tests/tree_contacts_solver.rs, on the float
  instantiation.  A sphere tree and a sphere body pushed apart conserve
  momentum to 1e-4 under PGS and Newton (:404-448).  The tilted tree box on
  a box body gives newt's single solver contact, at its position and depth;
  it pushes the tree up and the body down (:165-208).
- a reduction.  With the body replaced by a free-root tree of the same mass
  and inertia (a tree-tree contact, which newt's tree goldens pin), one
  Euler step from rest gives the same velocities to rounding in double.  The
  spheres exercise the linear rows and the tilted boxes the lever arms.
"""
import numpy as np
import pytest

from tooling_amd import abi, scenes

SOLVERS = [abi.SOLVER_PGS, abi.SOLVER_NEWTON]


def _rot(q, v):
    x, y, z, w = q
    u = np.array([x, y, z])
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def _one_step(oracle_lib, sc, precision):
    ow = oracle_lib.OracleWorlds(sc, 1, precision)
    q0, v0 = sc.initial_state(1, 0)
    ow.set_state(q0, v0)
    ow.step(1, threads=1)
    q, v = ow.state()
    ids, geo = ow.contacts(0)
    ow.close()
    return q[0], v[0], ids, geo


@pytest.mark.parametrize("solver", SOLVERS)
def test_tree_body_impulses_equal_and_opposite(oracle_lib, solver):
    sc = scenes.tree_body_spheres(solver)
    q, v, ids, _ = _one_step(oracle_lib, sc, 1)
    assert len(ids) == 1
    tree_p = v[3:6] * 1.0                 # the free root's body-frame velocity, identity orientation
    body_p = v[6:9] * 2.0
    assert np.abs(tree_p + body_p).max() < 1e-4, tree_p + body_p
    assert body_p[0] > 1e-3 and tree_p[0] < -1e-3


@pytest.mark.parametrize("solver", SOLVERS)
def test_tilted_tree_box_on_body(oracle_lib, solver):
    sc = scenes.tree_body_tilted_boxes(solver)
    q, v, ids, geo = _one_step(oracle_lib, sc, 1)
    assert len(ids) == 1
    np.testing.assert_allclose(geo[0, :3], [0.5, 0.4949751, 1.004975], atol=1e-5)
    assert abs(geo[0, 6] - 0.02250594) < 2e-5
    assert v[5] > 0.01          # tree root's body-frame z velocity
    assert v[8] < 0.0           # the body's world z velocity


@pytest.mark.parametrize("make", [scenes.tree_body_spheres, scenes.tree_body_tilted_boxes])
@pytest.mark.parametrize("solver", SOLVERS)
def test_body_side_equals_tree_side(oracle_lib, make, solver):
    kw = {"integrator": abi.INTEGRATOR_EULER} if make is scenes.tree_body_tilted_boxes else {}
    mixed = make(solver, **kw)
    trees = make(solver, as_tree=True, **kw)
    if make is scenes.tree_body_spheres:
        mixed.options["integrator"] = trees.options["integrator"] = abi.INTEGRATOR_EULER
    qm, vm, idm, gm = _one_step(oracle_lib, mixed, 0)
    qt, vt, idt, gt = _one_step(oracle_lib, trees, 0)
    assert len(idm) == len(idt) >= 1
    np.testing.assert_allclose(gm, gt, rtol=0, atol=1e-14)
    # the first tree alike
    np.testing.assert_allclose(vm[:6], vt[:6], rtol=0, atol=1e-12)
    # body (world linear, body angular) against the second tree's free root
    # (body angular, body linear); from rest the body-frame velocity turns
    # into the world frame by the start-of-step orientation.
    q_start = mixed.initial_state(1, 0)[0][0, 10:14]
    np.testing.assert_allclose(vm[6:9], _rot(q_start, vt[9:12]), rtol=0, atol=1e-12)
    np.testing.assert_allclose(vm[9:12], vt[6:9], rtol=0, atol=1e-12)
    assert np.abs(vm[6:12]).max() > 1e-3


def test_mixed_world_steps(oracle_lib):
    """The chain-and-boxes scene steps in both instantiations with all four
    contact kinds; momentum-free checks aside, float and double stay close
    over 100 steps."""
    sc = scenes.chain_and_boxes()
    kinds = set()
    out = []
    for precision in (0, 1):
        ow = oracle_lib.OracleWorlds(sc, 2, precision)
        q0, v0 = sc.initial_state(2, 3)
        ow.set_state(q0, v0)
        for s in range(100):
            ow.step(1, threads=2)
            if precision == 0:
                for w in range(2):
                    ids, _ = ow.contacts(w)
                    for a, b, _slot in ids:
                        ga, gb = sc.geoms[a], sc.geoms[b]
                        kind = lambda g: "tree" if g["tree"] >= 0 else "body" if g["body"] >= 0 else "static"
                        kinds.add(tuple(sorted((kind(ga), kind(gb)))))
        out.append(ow.state())
        ow.close()
    assert {("body", "tree"), ("body", "body"), ("body", "static"), ("static", "tree")} <= kinds, kinds
    assert np.isfinite(out[0][0]).all()
    assert np.abs(out[0][0] - out[1][0]).max() < 1e-2
