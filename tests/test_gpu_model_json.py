"""Models loaded from newt's JSON format, stepped on the GPU vs the CPU
oracle on the same seeded states: newt's own models/pendulum.json (RK4,
penalty mode, no contacts) and the quadruped written in the format
(tests/golden/models/quadruped18.json), and newt's models/stack.json (three
free boxes) under its own penalty solver, PGS and Newton.  EXACT bitwise,
FP64 1e-9 (free bodies: bitwise in every mode)."""
import os
import sys

import numpy as np
import pytest

from tooling_amd import abi, model_json, scenes
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu

MODELS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "models")
TOL_FP64 = 1e-9


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b, floor=1.0):
    den = np.maximum(np.abs(b).max(axis=-1), floor)
    return (np.abs(np.asarray(a) - np.asarray(b)).max(axis=-1) / den).max()


@pytest.fixture(scope="module")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


def _compare(oracle, sc, q0, v0, steps, precision):
    n = q0.shape[0]
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    step(m, d, steps)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    ow.step(steps, threads=8)
    (q, v), (qr, vr) = d.state(), ow.state()
    assert not d.faults().any()
    assert rel_err(q, qr) <= TOL_FP64 and rel_err(v, vr) <= TOL_FP64
    if precision == abi.PRECISION_FP64_EXACT:
        assert np.array_equal(q, qr) and np.array_equal(v, vr)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_newt_pendulum_json(oracle, precision):
    sc = model_json.load(os.path.join(MODELS, "pendulum.json"))
    rng = np.random.default_rng(3)
    q0 = rng.uniform(-1.0, 1.0, size=(64, 2))
    v0 = rng.normal(0.0, 0.5, size=(64, 2))
    say(f"pendulum.json precision {precision}: 64 worlds x 500 steps")
    _compare(oracle, sc, q0, v0, 500, precision)


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_quadruped_json(oracle, precision):
    sc = model_json.load(os.path.join(MODELS, "quadruped18.json"))
    q0, v0 = scenes.quadruped().initial_state(64, 13)
    say(f"quadruped18.json precision {precision}: 64 worlds x 200 steps")
    _compare(oracle, sc, q0, v0, 200, precision)


@pytest.mark.parametrize("solver", [abi.SOLVER_PENALTY, abi.SOLVER_PGS, abi.SOLVER_NEWTON])
def test_stack_json_free_bodies(oracle, solver):
    """newt's stack.json loads as three free boxes; under its own penalty
    solver, PGS and Newton the GPU matches the oracle bit for bit, with
    random velocity perturbations per world."""
    sc = model_json.load(os.path.join(MODELS, "stack.json"))
    if solver != abi.SOLVER_PENALTY:
        sc = scenes.with_options(sc, f"_s{solver}", solver=solver, iterations=30)
    q0, v0 = sc.initial_state(32, 0)
    v0 = v0 + np.random.default_rng(5).normal(0.0, 0.05, size=v0.shape)
    say(f"stack.json solver {solver}: 32 worlds x 400 steps")
    _compare(oracle, sc, q0, v0, 400, abi.PRECISION_FP64_EXACT)
