"""newt_batch: datas of different models (hinge chains of 2..64 links and the
quadruped: three register widths, two step variants) stepped by one launch
per kernel instantiation, replayed from a HIP graph.

- every world of the batch equals the same data stepped alone by newt_step,
  bit for bit, whatever the batch options (order, workgroup caps, graph,
  branches);
- FP64: within 1e-9 of the FP64 oracle on qpos, qvel and qacc over 100
  steps; mixed: qacc within 1e-5 of the oracle one step from the step-100
  state.
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Batch, Data, Model, step

pytestmark = pytest.mark.gpu

LENGTHS = (2, 5, 17, 24, 25, 31, 40, 64)
PER_MODEL = 6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, r):
    return float((np.abs(a - r).max(axis=1) / np.maximum(np.abs(r).max(axis=1), 1.0)).max())


def _models():
    out = []
    for n in LENGTHS:
        sc = scenes.hinge_chain(n)
        out.append((sc, Model(sc), *sc.initial_state(PER_MODEL, 100 + n)))
    sc = scenes.quadruped()
    q, v = sc.initial_state(PER_MODEL, 7)
    q[:, 2] = 0.26  # feet on the plane from the start
    out.append((sc, Model(sc), q, v))
    return out


def _datas(models, precision):
    ds = []
    for sc, m, q, v in models:
        d = Data(m, q.shape[0], precision)
        d.set_state(q, v)
        ds.append(d)
    return ds


def _snap(ds):
    return [(*d.state(), d.qacc(), d.contacts()[0], d.faults()) for d in ds]


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_MIXED])
def test_batch_equals_each_data_alone(precision):
    models = _models()
    alone = _datas(models, precision)
    for (sc, m, _, _), d in zip(models, alone):
        step(m, d, 60)
    ref = _snap(alone)
    for opts in (dict(), dict(order=0, use_graph=0, branches=0), dict(wpb_cap=(2, 1, 2), branches=0),
                 dict(order=0, wpb_cap=(1, 2, 1))):
        _say(f"batch precision {precision} options {opts}")
        ds = _datas(models, precision)
        b = Batch(ds, **opts)
        assert b.launches() == 3  # register widths 24 (chains of <= 24 links and the quadruped), 32, 64
        for _ in range(3):
            b.step(20)
        got = _snap(ds)
        b.close()
        for (sc, _, _, _), g, r in zip(models, got, ref):
            for k in range(5):
                np.testing.assert_array_equal(g[k], r[k], err_msg=f"{sc.name} field {k} options {opts}")
        for d in ds:
            d.close()
    for d in alone:
        d.close()


def test_batch_fp64_against_oracle_over_100_steps(oracle_lib):
    models = _models()
    ds = _datas(models, abi.PRECISION_FP64)
    b = Batch(ds)
    b.step(50)
    b.step(50)
    b.close()
    for (sc, m, q, v), d in zip(models, ds):
        _say(f"batch fp64 vs oracle: {sc.name}")
        ow = oracle_lib.OracleWorlds(sc, q.shape[0], 0)
        ow.set_state(q, v)
        ow.step(100, threads=q.shape[0])
        gq, gv = d.state()
        oq, ov = ow.state()
        assert _rel(gq, oq) <= 1e-9, sc.name
        assert _rel(gv, ov) <= 1e-9, sc.name
        assert _rel(d.qacc(), ow.qacc()) <= 1e-9, sc.name
        ow.close()
        d.close()


def test_batch_mixed_qacc_against_oracle(oracle_lib):
    models = _models()
    ds = _datas(models, abi.PRECISION_MIXED)
    b = Batch(ds)
    b.step(100)
    states = [d.state() for d in ds]
    b.step(1)
    b.close()
    touching = 0
    for (sc, m, _, _), d, (qs, vs) in zip(models, ds, states):
        ow = oracle_lib.OracleWorlds(sc, qs.shape[0], 0)
        ow.set_state(qs, vs)
        ow.step(1, threads=qs.shape[0])
        assert _rel(d.qacc(), ow.qacc()) <= 1e-5, sc.name
        touching += int((d.contacts()[0] > 0).sum())
        ow.close()
        d.close()
    assert touching > 0
