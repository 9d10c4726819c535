"""Box-box contacts on the GPU (box_box_full_manifold, newt/src/contact.rs:
569-1196: 15-axis SAT, face clipping, edge-edge) vs the CPU oracle, whose
box-box manifold is pinned by the solver_tree golden.  Scene: hinged boxes
between two static box walls, with box-plane contacts beside them and over
64 rows in many worlds.

EXACT: pair sets and contact (geom_a, geom_b, slot) lists identical at every
step and the state bitwise; FP64 within 1e-9 relative; mixed one-step qacc
within 1e-5.
"""
import sys

import numpy as np
import pytest

from tooling_amd import abi, scenes
from tooling_amd.newt import Data, Model, step

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-9
TOL_MIXED_QACC = 1e-5


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b, floor=1.0):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.maximum(np.abs(b).max(axis=-1), floor)
    return (np.abs(a - b).max(axis=-1) / den).max()


@pytest.fixture(scope="module")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_box_corridor_pairs_contacts_and_state(oracle, precision):
    sc = scenes.box_corridor(6)
    n = 32
    q0, v0 = sc.initial_state(n, 5)
    say(f"{sc.name} precision {precision}: {n} worlds x 150 single steps, pairs/contacts each step")
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    box_box = 0
    for s in range(150):
        step(m, d, 1)
        ow.step(1, threads=8)
        pc, pairs = d.pairs()
        cc, ids, geo = d.contacts()
        for w in range(n):
            pr = ow.pairs(w)
            assert pc[w] == len(pr) and np.array_equal(pairs[w, :pc[w]], pr), (s, w)
            ir, gr = ow.contacts(w)
            assert cc[w] == len(ir) and np.array_equal(ids[w, :cc[w]], ir), (s, w)
            if cc[w]:
                assert rel_err(geo[w, :cc[w]], gr) <= TOL_FP64
                box_box += sum(1 for a, b, _ in ir if sc.geoms[a]["shape"] == sc.geoms[b]["shape"] == abi.SHAPE_BOX)
    say(f"box-box contacts seen: {box_box}")
    assert box_box > 0
    assert not d.faults().any()
    q, v = d.state()
    qr, vr = ow.state()
    assert rel_err(q, qr) <= TOL_FP64 and rel_err(v, vr) <= TOL_FP64
    assert rel_err(d.qacc(), ow.qacc()) <= TOL_FP64
    if precision == abi.PRECISION_FP64_EXACT:
        assert np.array_equal(q, qr) and np.array_equal(v, vr)


def test_box_corridor_mixed_one_step_qacc(oracle):
    sc = scenes.box_corridor(6)
    n = 64
    q0, v0 = sc.initial_state(n, 9)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    m = Model(sc)
    worst = 0.0
    for t in (0, 50, 100, 100):
        ow.step(t, threads=8)
        qs, vs = ow.state()
        d = Data(m, n, abi.PRECISION_MIXED)
        d.set_state(qs, vs)
        step(m, d, 1)
        ref = oracle.OracleWorlds(sc, n, 0)
        ref.set_state(qs, vs)
        ref.step(1, threads=8)
        worst = max(worst, rel_err(d.qacc(), ref.qacc()))
        d.close()
        ref.close()
    say(f"box corridor mixed qacc relative error {worst:.3e}")
    assert worst <= TOL_MIXED_QACC
