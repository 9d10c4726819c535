"""The Newton contact solve on trees (the tree solver's Newton branch,
solver.rs:3161-3241, running the nonnegative Newton of newton.rs:275-556 on
H = A + diag(reg)) on the GPU's lane-per-world step, against the FP64 oracle,
whose restatement reproduces newt's solver_newton_tree.bin byte for byte
(test_oracle_golden.py).  Tree models under Newton step on the lane kernel
in every precision, in newt's operation order, so every case is bitwise:

- solver_newton_tree's scene (a free box with a hinged box below it, RK4,
  iterations 30, dropped onto a plane of friction 0.7) in 64 worlds with
  perturbed poses, 600 steps;
- the quadruped landing (Euler, 50 iterations): contact lists identical
  every 10th step, state and qacc bitwise after 200 steps;
- the same with frictionless feet: the four rows of a zero-friction contact
  are equal, H is only semidefinite and the solve adds newt's solver-only
  diagonal pivot;
- the box corridor (hinged boxes between box walls: box-box manifolds and
  more than 16 rows in a world);
- ranged joints and acceleration sensors under Newton are rejected at
  finalize (newt has no Newton joint-limit rows; the sensors' contact
  re-solve is built for PGS).
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Data, Model, step
from tooling_amd.scenes import Scene, geom, link, sensor

pytestmark = pytest.mark.gpu

PRECISIONS = [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64, abi.PRECISION_MIXED]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def newton_tree_scene():
    """newt/tests/solver_golden.rs's tree-contact scene under Newton, FP64."""
    plane = geom(abi.SHAPE_PLANE, friction=0.7, solref=(0.05, 1.5))
    root = link(-1, abi.JOINT_FREE, mass=1.2, inertia=(0.08, 0.0, 0.0, 0.0, 0.09, 0.0, 0.0, 0.0, 0.1))
    lower = link(0, abi.JOINT_HINGE, axis=(0.0, 1.0, 0.0), parent_pos=(0.0, 0.0, -0.3), child_pos=(0.0, 0.0, 0.22),
                 mass=0.75, inertia=(0.02, 0.0, 0.0, 0.0, 0.025, 0.0, 0.0, 0.0, 0.018))
    b0 = geom(abi.SHAPE_BOX, tree=0, link=0, size=(0.22, 0.17, 0.19), pos=(0.08, -0.06, 0.0),
              quat=scenes.quat_axis_angle((0.0, 1.0, 0.0), 0.2), friction=0.7)
    b1 = geom(abi.SHAPE_BOX, tree=0, link=1, size=(0.14, 0.12, 0.24), pos=(0.0, 0.0, -0.18),
              quat=scenes.quat_axis_angle((1.0, 0.0, 0.0), -0.15), friction=0.7)
    ro = scenes.quat_axis_angle((1.0, 0.4, -0.2), 0.35)

    def sampler(n, seed):
        r = np.random.default_rng(seed)
        qpos = np.zeros((n, 8))
        qpos[:, :3] = (0.13, -0.07, 0.78)
        qpos[:, :3] += r.uniform(-0.05, 0.05, size=(n, 3))
        for w in range(n):
            axis = (1.0, 0.4, -0.2) + r.normal(0.0, 0.1, size=3)
            qpos[w, 3:7] = scenes.quat_axis_angle(tuple(axis), 0.35 + r.uniform(-0.2, 0.2))
        qpos[0, 3:7] = ro
        qpos[:, 7] = 0.37 + r.uniform(-0.3, 0.3, size=n)
        qvel = np.zeros((n, 7))
        qvel[:, 6] = -0.25
        qvel[:, :6] += r.normal(0.0, 0.3, size=(n, 6))
        return qpos, qvel

    opts = dict(dt=0.005, gravity=(0.0, 0.0, -9.81), integrator=abi.INTEGRATOR_RK4, solver=abi.SOLVER_NEWTON,
                iterations=30)
    return Scene("newton_tree", opts, [root, lower], [plane, b0, b1], 1, sampler,
                 "free box over a hinged box on a plane, Newton, RK4")


def _run(oracle_lib, sc, n, steps, precision, seed=3, contacts_every=0):
    q0, v0 = sc.initial_state(n, seed)
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle_lib.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    rows = 0
    if contacts_every:
        for s in range(steps):
            step(m, d, 1)
            ow.step(1, threads=8)
            if s % contacts_every == contacts_every - 1:
                counts, gids, _ = d.contacts()
                for w in range(n):
                    oids, _ = ow.contacts(w)
                    assert [tuple(x) for x in gids[w, :counts[w]]] == [tuple(x) for x in oids], (s, w)
                    rows = max(rows, 4 * int(counts[w]))
    else:
        for chunk in (1, steps // 2 - 1, steps - steps // 2):
            step(m, d, chunk)
            ow.step(chunk, threads=8)
    return d, ow, rows


def _assert_bitwise(d, ow):
    assert not d.faults().any()
    q, v = d.state()
    oq, ov = ow.state()
    np.testing.assert_array_equal(q, oq)
    np.testing.assert_array_equal(v, ov)
    np.testing.assert_array_equal(d.qacc(), ow.qacc())


@pytest.mark.parametrize("precision", PRECISIONS)
def test_newton_tree_scene(oracle_lib, precision):
    sc = newton_tree_scene()
    d, ow, _ = _run(oracle_lib, sc, 64, 600, precision)
    _say(f"newton tree scene precision {precision}: contacts now {int(d.contacts()[0].sum())}")
    _assert_bitwise(d, ow)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_MIXED])
@pytest.mark.parametrize("friction", [0.8, 0.0])
def test_quadruped_landing_under_newton(oracle_lib, precision, friction):
    sc = scenes.with_options(scenes.quadruped(friction=friction), "_newton", solver=abi.SOLVER_NEWTON)
    d, ow, rows = _run(oracle_lib, sc, 32, 200, precision, contacts_every=10)
    _say(f"quadruped newton friction {friction} precision {precision}: most rows {rows}")
    assert rows > 0
    _assert_bitwise(d, ow)


def test_box_corridor_under_newton(oracle_lib):
    sc = scenes.with_options(scenes.box_corridor(6), "_newton", solver=abi.SOLVER_NEWTON)
    d, ow, rows = _run(oracle_lib, sc, 16, 150, abi.PRECISION_FP64, contacts_every=10)
    _say(f"box corridor newton: most rows {rows}")
    assert rows > 16
    _assert_bitwise(d, ow)


def test_newton_rejections():
    q = scenes.quadruped()
    ranged = [dict(l) for l in q.links]
    ranged[3] = dict(ranged[3], range=(-0.5, 0.5))
    sc = Scene("ranged_newton", dict(q.options, solver=abi.SOLVER_NEWTON), ranged, q.geoms, 0, q.sampler, "")
    with pytest.raises(abi.NewtError) as e:
        Model(sc)
    assert e.value.status == abi.ERR_UNSUPPORTED
    nt = scenes.with_options(q, "_newton", solver=abi.SOLVER_NEWTON)
    acc = Scene("acc_newton", nt.options, nt.links, nt.geoms, 0, nt.sampler, "",
                sensors=[sensor(abi.SENSOR_ACCELEROMETER, link=0)])
    with pytest.raises(abi.NewtError) as e:
        Model(acc)
    assert e.value.status == abi.ERR_UNSUPPORTED
    kin = Scene("kin_newton", nt.options, nt.links, nt.geoms, 0, nt.sampler, "",
                sensors=[sensor(abi.SENSOR_FRAME_POS, link=0), sensor(abi.SENSOR_JOINT_POS, link=2)])
    Model(kin).close()
