"""Whole-step pins against MuJoCo: newt's differential suite
(newt/tests/differential.rs) restated on the float instantiation of the
oracle.  Each scenario's MJCF (tests/golden/mujoco/) loads through
tooling_amd.mjcf, starts from scenarios.json's initial state, steps n_steps
and is compared at every stride with MuJoCo 3.11's recorded qpos / qvel
(NEWTDIF1 fixtures) within newt's own per-scenario bounds:

- ballistic and tumble (free bodies, RK4 and Euler), double_pendulum (a
  fixed-root chain, RK4 and Euler, 5e-7), floating_base (a free root with a
  hinge: every slot of the free-root remap, 6e-6);
- sphere_drop, its stiff and soft solref companions, box_stack (free-body
  PGS, RK4 and Euler; Newton Euler);
- joint_limit_swing (tree PGS limit rows, RK4);
- tree_chain_contact (a falling free-root hinge chain on a plane: tree
  contacts under Euler, with PGS and with Newton): the tree Euler + PGS and
  Newton steps as a whole.
"""
import numpy as np
import pytest

import differential as D
from oracle import pyoracle


@pytest.fixture(scope="module", autouse=True)
def _lib():
    pyoracle.lib()


@pytest.mark.parametrize("case", D.CASES, ids=D.case_id)
def test_mujoco_differential(case):
    name, suffix = case[0], case[1]
    sc = D.scene(case)
    stride, n_steps, fq, fv = D.read_fixture(f"{D.MJ}/{name}{suffix}.bin")
    s = D.spec(name)
    assert (stride, n_steps) == (s["stride"], s["n_steps"])
    q0, v0 = D.initial_state(case, sc)
    ow = pyoracle.OracleWorlds(sc, 1, 1)
    ow.set_state(q0[None], v0[None])
    samples = []
    for k in range(n_steps // stride + 1):
        if k:
            ow.step(stride, threads=1)
        q, v = ow.state()
        samples.append(D.to_mujoco(sc, q[0], v[0]))
    ow.close()
    dq, dv = D.divergence(sc, fq, fv, samples)
    print(f"{name}{suffix}: qpos {dq:.3e} (bound {case[4]:.1e}), qvel {dv:.3e} (bound {case[5]:.1e})")
    assert dq <= case[4] and dv <= case[5]
