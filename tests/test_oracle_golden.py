"""The oracle against newt's byte-exact golden trajectories.

Each fixture under tests/golden/ is a file from newt/tests/goldens/ (raw LE
f32 state snapshots written by the reference's own tests).  The oracle's
float instantiation rebuilds the same scene and must reproduce every byte:
  joints_chain_and_floating  ABA + FK + RK4, fixed chain and free root
                             (newt/tests/joints_golden.rs)
  joints_mixed               hinge + slide + ball + free root, RK4
                             (newt/tests/joints_mixed_golden.rs)
  dynamics_mass_bias         CRB mass matrix and RNE bias
                             (newt/tests/dynamics_golden.rs)
  solver_tree                broadphase, box-plane + box-box manifold, tree
                             contact PGS (Delassus, diagApprox, SolImp/SolRef),
                             RK4 (newt/tests/solver_golden.rs:93-144)
  tumbling_3_body            free-body RK4: Euler's equations, quaternion
                             derivative, renormalize (newt/tests/golden.rs)
  integrator_euler / _rk4 /  one free body under each integrator, incl.
  integrator_implicitfast    integrate_body_angular_velocity
                             (newt/tests/integrators_golden.rs)
  solver_stack               free-body PGS: box-plane and box-box contacts,
                             rows, regularization, pyramidal facet pairs, body
                             deltas, wrench recovery, RK4 ZOH
                             (newt/tests/solver_golden.rs:27-46)
  solver_incline             free-body PGS friction under tilted gravity
                             (newt/tests/solver_golden.rs:48-79)
  solver_newton_stack /      the same scenes under the Newton solver: dense
  solver_newton_incline      Delassus from unit impulses, Cholesky, Newton
                             directions and the exact projected line search
                             (newton.rs:275-556; solver_golden.rs:81-91)
  solver_newton_tree         the tree-contact scene of solver_tree under the
                             Newton solver: H = A + diag(R) of the tree rows,
                             every row projected onto f >= 0
                             (solver.rs:3161-3241; solver_golden.rs:265-269)
  sensors_battery            sensordata of a pendulum tree and a tumbling
                             free body in one world after steps 1, 100 and
                             500: joint position/velocity, frame position and
                             orientation on a link, gyro and accelerometer on
                             a body (newt/tests/sensors_battery.rs:1056-1176)
"""
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["joints_chain_and_floating", "joints_mixed", "dynamics_mass_bias", "solver_tree",
                                  "tumbling_3_body", "integrator_euler", "integrator_rk4",
                                  "integrator_implicitfast", "solver_stack", "solver_incline",
                                  "solver_newton_stack", "solver_newton_incline", "solver_newton_tree", "sensors_battery"])
def test_golden_bytes(oracle_lib, name):
    expected = open(os.path.join(GOLDEN, name + ".bin"), "rb").read()
    actual = oracle_lib.golden_bytes(name)
    assert len(actual) == len(expected)
    if actual != expected:
        first = next(i for i, (a, b) in enumerate(zip(actual, expected)) if a != b)
        pytest.fail(f"{name}: first differing byte at offset {first}")
