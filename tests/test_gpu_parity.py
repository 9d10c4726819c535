"""HIP path vs the CPU oracle on the same seeded inputs.

FP64 step: pair sets and contact (geom_a, geom_b, slot) lists identical to
the oracle at every step; qpos, qvel and qacc within 1e-9 relative
(‖Δ‖∞ / max(‖ref‖∞, 1), max over worlds) over 100+ steps.  Mixed step: one
step from the oracle's state gives qacc within 1e-5 relative, at states from
free fall to lying flat on the plane (the worst-conditioned contact sets).
For scale: newt's own f32 step is ~2e-3 from the FP64 oracle on the same
states.
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


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64, abi.PRECISION_FP64_EXACT])
def test_pendulum_fp64_100_steps(oracle, precision):
    sc = scenes.pendulum()
    n = 256
    q0, v0 = sc.initial_state(n, 7)
    say(f"pendulum3 precision {precision}: 256 worlds x 100 steps")
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    step(m, d, 100)
    q, v = d.state()
    a = d.qacc()
    assert not d.faults().any()
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    ow.step(100, threads=8)
    qr, vr = ow.state()
    ar = ow.qacc()
    assert rel_err(q, qr) <= TOL_FP64
    assert rel_err(v, vr) <= TOL_FP64
    assert rel_err(a, ar) <= TOL_FP64
    if precision == abi.PRECISION_FP64_EXACT:
        assert np.array_equal(q, qr) and np.array_equal(v, vr) and np.array_equal(a, ar)


def test_pendulum_fp64_1000_steps_one_launch(oracle):
    sc = scenes.pendulum()
    n = 64
    q0, v0 = sc.initial_state(n, 11)
    say("pendulum3 fp64: 64 worlds x 1000 steps")
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(q0, v0)
    step(m, d, 1000)
    q, v = d.state()
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    ow.step(1000, threads=8)
    qr, vr = ow.state()
    # 1000 steps of a chaotic triple pendulum: report, bound loosely.
    assert rel_err(q, qr) <= 1e-6


@pytest.mark.parametrize("precision", [abi.PRECISION_FP64_EXACT, abi.PRECISION_FP64])
def test_quadruped_fp64_pairs_contacts_and_state(oracle, precision):
    sc = scenes.quadruped()
    n = 64
    q0, v0 = sc.initial_state(n, 3)
    say(f"quadruped18 precision {precision}: 64 worlds x 250 single steps, pairs/contacts each step")
    m = Model(sc)
    d = Data(m, n, precision)
    d.set_state(q0, v0)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    touched = 0
    for s in range(250):
        step(m, d, 1)
        ow.step(1, threads=8)
        pc, pairs = d.pairs()
        cc, ids, geo = d.contacts()
        for w in range(n):
            pr = ow.pairs(w)
            assert pc[w] == len(pr)
            assert np.array_equal(pairs[w, :pc[w]], pr), (s, w)
            ir, gr = ow.contacts(w)
            assert cc[w] == len(ir), (s, w)
            assert np.array_equal(ids[w, :cc[w]], ir), (s, w)
            if cc[w]:
                touched += 1
                assert rel_err(geo[w, :cc[w]], gr) <= TOL_FP64
    assert touched > 0
    assert not d.faults().any()
    q, v = d.state()
    qr, vr = ow.state()
    assert rel_err(q, qr) <= TOL_FP64
    assert rel_err(v, vr) <= TOL_FP64
    assert rel_err(d.qacc(), ow.qacc()) <= TOL_FP64
    if precision == abi.PRECISION_FP64_EXACT:
        assert np.array_equal(q, qr) and np.array_equal(v, vr)


def test_quadruped_fp64_steady_state_100_steps(oracle):
    """100 steps from a contact-rich state (robots lying on the plane)."""
    sc = scenes.quadruped()
    n = 128
    q0, v0 = sc.initial_state(n, 9)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    ow.step(400, threads=8)
    qs, vs = ow.state()
    say("quadruped18 fp64: 128 worlds x 100 steps from step 400")
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(qs, vs)
    step(m, d, 100)
    ow.step(100, threads=8)
    q, v = d.state()
    qr, vr = ow.state()
    assert not d.faults().any()
    assert rel_err(q, qr) <= TOL_FP64
    assert rel_err(v, vr) <= TOL_FP64
    assert rel_err(d.qacc(), ow.qacc()) <= TOL_FP64


def test_quadruped_mixed_one_step_qacc(oracle):
    sc = scenes.quadruped()
    n = 256
    q0, v0 = sc.initial_state(n, 5)
    ow = oracle.OracleWorlds(sc, n, 0)
    ow.set_state(q0, v0)
    say("quadruped18 mixed: 256 worlds, qacc of one step from oracle states")
    m = Model(sc)
    worst = 0.0
    for t in (0, 60, 60, 60, 60, 60, 100, 100):  # free fall, landing, lying flat
        ow.step(t, threads=8)
        qs, vs = ow.state()
        d = Data(m, n, abi.PRECISION_MIXED)
        d.set_state(qs, vs)
        step(m, d, 1)
        ref = oracle.OracleWorlds(sc, n, 0)
        ref.set_state(qs, vs)
        ref.step(1, threads=8)
        err = rel_err(d.qacc(), ref.qacc())
        worst = max(worst, err)
        d.close()
        ref.close()
    say(f"mixed qacc relative error {worst:.3e}")
    assert worst <= TOL_MIXED_QACC


def _settled(oracle, sc, n, seed, steps):
    ow = oracle.OracleWorlds(sc, n, 0)
    q0, v0 = sc.initial_state(n, seed)
    ow.set_state(q0, v0)
    ow.step(steps, threads=8)
    return ow


@pytest.mark.parametrize("n_boxes", [6, 20])
def test_box_chain_fp64_contact_sets_over_64_rows(oracle, n_boxes):
    """Up to 4*n_boxes contacts: rows past the 64 a wavefront keeps in
    registers take the global-memory row path; 20 boxes (nv 25) also take
    the wider register-array instantiation."""
    sc = scenes.box_chain(n_boxes)
    n = 32
    ow = _settled(oracle, sc, n, 4, 60)
    qs, vs = ow.state()
    say(f"{sc.name} fp64: {n} worlds x 40 steps from a settled state")
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(qs, vs)
    most = 0
    for _ in range(4):
        step(m, d, 10)
        ow.step(10, threads=8)
        cc, ids, _geo = d.contacts()
        for w in range(n):
            ir, _gr = ow.contacts(w)
            assert cc[w] == len(ir)
            assert np.array_equal(ids[w, :cc[w]], ir)
        most = max(most, int(cc.max()))
    assert most > 16, "no world reached more than 64 rows"
    assert not d.faults().any()
    q, v = d.state()
    qr, vr = ow.state()
    assert rel_err(q, qr) <= TOL_FP64
    assert rel_err(v, vr) <= TOL_FP64
    assert rel_err(d.qacc(), ow.qacc()) <= TOL_FP64


def test_box_chain_mixed_one_step_qacc(oracle):
    sc = scenes.box_chain(6)
    n = 64
    ow = _settled(oracle, sc, n, 6, 80)
    qs, vs = ow.state()
    say(f"{sc.name} mixed: {n} worlds, qacc of one step from settled states")
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_MIXED)
    d.set_state(qs, vs)
    step(m, d, 1)
    ow.step(1, threads=8)
    err = rel_err(d.qacc(), ow.qacc())
    say(f"mixed qacc relative error {err:.3e}")
    assert err <= TOL_MIXED_QACC


@pytest.mark.parametrize("n_links", [40, 64])
def test_hinge_chain_fp64_wide(oracle, n_links):
    """Chains of 40 and 64 hinges: the 64-wide register instantiation, LDS
    above 64 KiB per workgroup, long level schedules."""
    sc = scenes.hinge_chain(n_links)
    n = 16
    ow = _settled(oracle, sc, n, 8, 100)
    qs, vs = ow.state()
    say(f"{sc.name} fp64: {n} worlds x 60 steps from step 100")
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_FP64)
    d.set_state(qs, vs)
    touched = 0
    for _ in range(6):
        step(m, d, 10)
        ow.step(10, threads=8)
        cc, ids, _geo = d.contacts()
        for w in range(n):
            ir, _gr = ow.contacts(w)
            assert cc[w] == len(ir)
            assert np.array_equal(ids[w, :cc[w]], ir)
            touched += int(cc[w] > 0)
    assert touched > 0
    assert not d.faults().any()
    q, v = d.state()
    qr, vr = ow.state()
    assert rel_err(q, qr) <= TOL_FP64
    assert rel_err(v, vr) <= TOL_FP64


def test_hinge_chain_mixed_one_step_qacc(oracle):
    sc = scenes.hinge_chain(40)
    n = 32
    ow = _settled(oracle, sc, n, 9, 200)
    qs, vs = ow.state()
    say(f"{sc.name} mixed: {n} worlds, qacc of one step from step 200")
    m = Model(sc)
    d = Data(m, n, abi.PRECISION_MIXED)
    d.set_state(qs, vs)
    step(m, d, 1)
    ow.step(1, threads=8)
    err = rel_err(d.qacc(), ow.qacc())
    say(f"mixed qacc relative error {err:.3e}")
    assert err <= TOL_MIXED_QACC
