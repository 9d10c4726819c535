"""Penalty contacts on trees (newt's default solver: tree_wrenches_from_pairs
and apply_tree_contact_wrench, newt/src/world.rs:2570-2741, sampled at every
RK4 stage or once per Euler step) on the GPU vs the FP64 oracle, same seeded
inputs.  These models run the lane-per-world FP64 step in every precision
mode, so state, qacc and contacts must equal the oracle bit for bit.  The
oracle's tree penalty path is pinned by closed forms
(tests/test_oracle_penalty_tree.py).  Cases: the quadruped (capsule- and
box-plane pairs) under Euler, ImplicitFast and RK4; with penalty joint
limits; the self-colliding hinge coil (sphere and capsule pairs between two
links of one tree: the wrench on each side with its own arm); the box
corridor (box-box pairs between links and against static boxes through the
legacy vertex-face manifold, contact.rs:587-714).
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def _run(oracle, sc, n, seed, steps, chunks, precision):
    q0, v0 = sc.initial_state(n, seed)
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    for _ in range(chunks):
        step(m, d, steps // chunks)
        ow.step(steps // chunks, threads=8)
    f = d.faults()
    assert not f.any(), np.unique(f)
    for name, x, y in (("qpos", d.state()[0], ow.state()[0]), ("qvel", d.state()[1], ow.state()[1]),
                       ("qacc", d.qacc(), ow.qacc())):
        if not np.array_equal(x, y):
            bad = np.argwhere(x != y)
            pytest.fail(f"{name} differs in {len(bad)} entries, first {bad[0]} (max abs {np.abs(x - y).max():.3e})")
    nc, ids_g, geo_g = d.contacts()
    for w in range(n):
        ids, geo = ow.contacts(w)
        assert nc[w] == len(ids), (w, nc[w], len(ids))
        assert np.array_equal(ids_g[w, :nc[w]], ids), w
        assert np.array_equal(geo_g[w, :nc[w]], geo), w
    say(f"{sc.name}: contacts per world at the end {nc.min()}..{nc.max()}")
    d.close()
    m.close()
    ow.close()
    return nc


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_EULER, abi.INTEGRATOR_IMPLICITFAST, abi.INTEGRATOR_RK4])
@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_quadruped_penalty_contacts(oracle, integrator, precision):
    sc = scenes.with_options(scenes.quadruped(), f"_pen_i{integrator}", solver=abi.SOLVER_PENALTY,
                             integrator=integrator)
    say(f"{sc.name} precision {precision}: 64 worlds x 300 steps")
    nc = _run(oracle, sc, 64, 23, 300, 3, precision)
    assert nc.max() >= 4


def test_quadruped_penalty_contacts_and_limits(oracle):
    """Penalty contacts and penalty joint limits (stiffness 100, damping 0.5:
    JointLimit::DEFAULT's 1000 / 60 is unstable on these light links at
    dt 2 ms, in newt as in the oracle) in one step."""
    base = scenes.quadruped()
    links = [dict(l, range=(-0.15, 0.15) if l["joint"] == abi.JOINT_HINGE else None, limit_stiffness=100.0,
                  limit_damping=0.5) for l in base.links]
    sc = scenes.Scene(base.name + "_pen_ranged", dict(base.options, solver=abi.SOLVER_PENALTY), links, base.geoms,
                      base.self_collide, base.sampler, base.description)
    say(f"{sc.name}: 64 worlds x 300 steps")
    _run(oracle, sc, 64, 29, 300, 3, abi.PRECISION_FP64)
    ow = oracle.OracleWorlds(sc, 4, 0)
    ow.set_state(*sc.initial_state(4, 29))
    ow.step(300)
    q = ow.state()[0]
    ow.close()
    assert np.isfinite(q).all() and (np.abs(q[:, 7:]) > 0.15).any()   # springs engaged


@pytest.mark.parametrize("integrator", [abi.INTEGRATOR_EULER, abi.INTEGRATOR_RK4])
def test_coil_self_collision_penalty(oracle, integrator):
    sc = scenes.with_options(scenes.hinge_coil(), f"_pen_i{integrator}", solver=abi.SOLVER_PENALTY,
                             integrator=integrator)
    say(f"{sc.name}: 16 worlds x 300 steps")
    _run(oracle, sc, 16, 3, 300, 3, abi.PRECISION_FP64)


def test_box_corridor_penalty_legacy_manifold(oracle):
    """The corridor's boxes with solref time constant 0.1 (the default 0.02
    is unstable for these light boxes at dt 2 ms under penalty contacts, in
    the oracle as in newt's arithmetic), RK4; contacts compared every 10
    steps and link-link box-box contacts required."""
    base = scenes.box_corridor(6)
    geoms = [dict(g, solref=(0.1, 1.0)) for g in base.geoms]
    sc = scenes.Scene(base.name + "_pen", dict(base.options, solver=abi.SOLVER_PENALTY,
                                               integrator=abi.INTEGRATOR_RK4), base.links, geoms,
                      base.self_collide, base.sampler, base.description)
    n = 16
    q0, v0 = sc.initial_state(n, 3)
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    link_box = 0
    for _ in range(20):
        step(m, d, 10)
        ow.step(10, threads=8)
        nc, ids_g, geo_g = d.contacts()
        for w in range(n):
            ids, geo = ow.contacts(w)
            assert nc[w] == len(ids) and np.array_equal(ids_g[w, :nc[w]], ids), w
            assert np.array_equal(geo_g[w, :nc[w]], geo), w
            link_box += sum(1 for a, b, _ in ids if sc.geoms[a]["shape"] == sc.geoms[b]["shape"] == abi.SHAPE_BOX
                            and sc.geoms[a]["link"] >= 0 and sc.geoms[b]["link"] >= 0)
    assert not d.faults().any()
    np.testing.assert_array_equal(d.state()[0], ow.state()[0])
    np.testing.assert_array_equal(d.state()[1], ow.state()[1])
    say(f"{sc.name}: link-link box-box contacts seen {link_box}")
    assert link_box > 0
