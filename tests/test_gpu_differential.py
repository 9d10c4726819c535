"""newt's MuJoCo differential scenarios (newt/tests/differential.rs) on the
GPU: each MJCF scenario (tooling_amd.mjcf) steps on the device from
scenarios.json's initial state; at every stride the state must equal the
FP64 oracle's bit for bit under FP64_EXACT, and in every precision stay
within newt's own bound of MuJoCo 3.11's recorded trajectory (the bounds
the float oracle meets, tests/test_oracle_differential.py)."""
import sys

import numpy as np
import pytest

import differential as D
from tooling_amd import abi
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64, abi.PRECISION_MIXED])
@pytest.mark.parametrize("case", D.CASES, ids=D.case_id)
def test_gpu_mujoco_differential(oracle_lib, case, precision):
    name, suffix = case[0], case[1]
    sc = D.scene(case)
    stride, n_steps, fq, fv = D.read_fixture(f"{D.MJ}/{name}{suffix}.bin")
    q0, v0 = D.initial_state(case, sc)
    m = Model(sc)
    d = Data(m, 1, precision)
    d.set_state(q0[None], v0[None])
    ow = oracle_lib.OracleWorlds(sc, 1, 0)
    ow.set_state(q0[None], v0[None])
    samples = []
    for k in range(n_steps // stride + 1):
        if k:
            step(m, d, stride)
            ow.step(stride, threads=1)
        q, v = d.state()
        if precision == abi.PRECISION_FP64_EXACT:
            oq, ov = ow.state()
            np.testing.assert_array_equal(q, oq)
            np.testing.assert_array_equal(v, ov)
        samples.append(D.to_mujoco(sc, q[0], v[0]))
    assert not d.faults().any()
    dq, dv = D.divergence(sc, fq, fv, samples)
    print(f"{name}{suffix} precision {precision}: qpos {dq:.3e} (bound {case[4]:.1e}), qvel {dv:.3e} "
          f"(bound {case[5]:.1e})", flush=True)
    sys.stdout.flush()
    assert dq <= case[4] and dv <= case[5]
