"""The oracle's counting scalar changes nothing but the counts: stepping the
quadruped (contacts, PGS) and the pendulum with it gives the FP64 oracle's
states bit for bit, and every operation class is counted."""
import numpy as np

from tooling_amd import scenes


def test_counting_scalar_is_the_fp64_arithmetic(oracle_lib):
    from oracle import pyoracle
    for name in ("quadruped18", "pendulum3"):
        sc = scenes.SCENES[name]()
        q0, v0 = sc.initial_state(4, 11)
        a = pyoracle.OracleWorlds(sc, 4, 0)
        b = pyoracle.OracleWorlds(sc, 4, 2)
        a.set_state(q0, v0)
        b.set_state(q0, v0)
        a.step(120)
        b.flops(reset=True)
        b.step(120)
        counts = b.flops(reset=True)
        qa, va = a.state()
        qb, vb = b.state()
        assert np.array_equal(qa, qb) and np.array_equal(va, vb), name
        assert np.array_equal(a.qacc(), b.qacc()), name
        add, mul, div, sqrt = counts
        assert add > 0 and mul > 0 and div > 0 and sqrt > 0, name
