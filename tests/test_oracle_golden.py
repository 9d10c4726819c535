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
"""
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["joints_chain_and_floating", "joints_mixed", "dynamics_mass_bias", "solver_tree"])
def test_golden_bytes(oracle_lib, name):
    expected = open(os.path.join(GOLDEN, name + ".bin"), "rb").read()
    actual = oracle_lib.golden_bytes(name)
    assert len(actual) == len(expected)
    if actual != expected:
        first = next(i for i, (a, b) in enumerate(zip(actual, expected)) if a != b)
        pytest.fail(f"{name}: first differing byte at offset {first}")
