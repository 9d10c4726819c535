"""World::penetration_slop on the GPU (append_contacts, newt/src/world.rs:
2517-2528): every narrow phase raises the contact's force-free gap to the
slop.  Against the FP64 oracle, whose slop test_oracle_slop.py pins:

- the quadruped with a 2 mm slop, 64 worlds x 250 steps: contacts equal
  every step; FP64_EXACT bitwise, the FP64 wavefront step within 1e-9;
- newt's golden box stack with a 5 mm slop under PGS, 32 worlds x 300
  steps: bitwise on the lane step (FP64_EXACT) and the wavefront step
  (FP64);
- the 64-box pile under Newton with a 1 mm slop, 16 worlds x 100 steps:
  FP64_EXACT bitwise, FP64 within 1e-9.
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, r):
    return float((np.abs(a - r).max(axis=1) / np.maximum(np.abs(r).max(axis=1), 1.0)).max())


def _run(oracle_lib, sc, n, steps, precision, seed, bitwise, check_contacts=False):
    _say(f"{sc.name} slop {sc.options['slop']} precision {precision}: {n} worlds x {steps} steps")
    q0, v0 = sc.initial_state(n, seed)
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle_lib.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    touched = 0
    for s in range(steps):
        step(m, d, 1)
        ow.step(1, threads=8)
        if check_contacts:
            cc, ids, _ = d.contacts()
            for w in range(n):
                ir, _ = ow.contacts(w)
                assert np.array_equal(ids[w, :cc[w]], ir), (s, w)
                touched += int(cc[w] > 0)
    assert not d.faults().any()
    got = d.state() + (d.qacc(),)
    ref = ow.state() + (ow.qacc(),)
    d.close()
    ow.close()
    for name, x, y in zip(("qpos", "qvel", "qacc"), got, ref):
        if bitwise:
            np.testing.assert_array_equal(x, y, err_msg=name)
        else:
            assert _rel(x, y) <= 1e-9, (name, _rel(x, y))
    return touched


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_quadruped_with_slop(oracle_lib, precision):
    sc = scenes.with_options(scenes.quadruped(), "_slop", slop=0.002)
    touched = _run(oracle_lib, sc, 64, 250, precision, 3, precision == abi.PRECISION_FP64_EXACT, True)
    assert touched > 0


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_box_stack_with_slop(oracle_lib, precision):
    sc = scenes.with_options(scenes.box_stack(integrator=abi.INTEGRATOR_EULER), "_slop", slop=0.005)
    _run(oracle_lib, sc, 32, 300, precision, 7, True, True)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_newton_box_pile_with_slop(oracle_lib, precision):
    sc = scenes.with_options(scenes.box_pile(), "_newton_slop", solver=abi.SOLVER_NEWTON, slop=0.001)
    _run(oracle_lib, sc, 16, 100, precision, 11, precision == abi.PRECISION_FP64_EXACT)
