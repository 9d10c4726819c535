"""Continuous-batching server: slot allocator and request mix on the CPU;
on the GPU, every request served in a mix (shared launches, reused slots,
admission waiting for slots, horizons split into chunks) returns the same
state, bit for bit, as the same request stepped alone, and the FP64 results
stay within 1e-9 of the oracle."""
import sys

import numpy as np
import pytest

from tooling_amd import abi, serve
from tooling_amd.newt import Data, Model, step


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_slot_ranges_first_fit_and_coalesce():
    s = serve.SlotRanges(100)
    a = s.alloc(30)
    b = s.alloc(30)
    c = s.alloc(30)
    assert (a, b, c) == (0, 30, 60)
    assert s.alloc(20) == -1 and s.fits(10) and not s.fits(11)
    s.release(b, 30)
    assert s.alloc(31) == -1
    assert s.alloc(25) == 30           # first fit reuses the hole
    s.release(30, 25)
    s.release(a, 30)
    assert s.free == [(0, 60), (90, 10)]
    s.release(c, 30)
    assert s.free == [(0, 100)] and s.n_free == 100


def test_slot_ranges_random_against_bitmap():
    rng = np.random.default_rng(3)
    s = serve.SlotRanges(257)
    used = np.zeros(257, bool)
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.5:
            f, n = live.pop(int(rng.integers(len(live))))
            assert used[f:f + n].all()
            used[f:f + n] = False
            s.release(f, n)
        else:
            n = int(rng.integers(1, 40))
            f = s.alloc(n)
            # first fit: the lowest start of a free run of length n
            runs = [i for i in range(257 - n + 1) if not used[i:i + n].any()]
            assert f == (runs[0] if runs else -1)
            if f >= 0:
                used[f:f + n] = True
                live.append((f, n))
        assert s.n_free == int((~used).sum())
        assert all(s.free[i][0] + s.free[i][1] < s.free[i + 1][0] for i in range(len(s.free) - 1))


def test_rollout_mix_ranges():
    reqs = serve.rollout_mix(300, seed=5)
    assert len(reqs) == 300
    arr = np.array([r.arrival for r in reqs])
    assert (np.diff(arr) > 0).all()
    assert 0.9 < len(reqs) / arr[-1] / 200.0 < 1.1          # Poisson rate 200/s
    assert all(1 <= r.n_worlds <= 512 and 100 <= r.horizon <= 2000 for r in reqs)
    keys = {r.key for r in reqs}
    assert {"pendulum3", "quadruped18"} <= keys
    chains = [int(k[5:]) for k in keys if k.startswith("chain")]
    assert chains and min(chains) >= 2 and max(chains) <= 64
    for r in reqs[:20]:
        sc = serve.scene_for(r.key)
        assert r.qpos.shape == (r.n_worlds, sc.nq) and r.qvel.shape == (r.n_worlds, sc.nv)


def _mix():
    """Eight requests over three models; the quadruped pool (48 slots) cannot
    hold its three requests at once, so the last waits and reuses slots."""
    spec = [("quadruped18", 30, 90, 0.0), ("pendulum3", 17, 70, 0.0), ("chain8", 9, 55, 0.0),
            ("quadruped18", 12, 40, 0.002), ("quadruped18", 25, 33, 0.003), ("pendulum3", 40, 23, 0.004),
            ("chain8", 20, 64, 0.005), ("pendulum3", 1, 100, 0.006)]
    reqs = []
    for i, (key, n, h, t) in enumerate(spec):
        q, v = serve.scene_for(key).initial_state(n, 100 + i)
        reqs.append(serve.Request(key, np.asarray(q, float), np.asarray(v, float), h, t))
    return reqs


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [abi.PRECISION_MIXED, abi.PRECISION_FP64])
def test_served_equals_alone(precision):
    reqs = _mix()
    pools = {"quadruped18": (serve.scene_for("quadruped18"), 48), "pendulum3": (serve.scene_for("pendulum3"), 64),
             "chain8": (serve.scene_for("chain8"), 32)}
    say(f"serve mix precision {precision}: {len(reqs)} requests, chunk 16")
    srv = serve.Server(pools, precision=precision, chunk=16, n_streams=4, hbm_budget=4 << 30)
    try:
        srv.run(reqs, time_limit=120.0)
        say(f"served: {srv.launches} launches, {srv.world_steps} world-steps")
        models = {k: p.model for k, p in srv.pools.items()}
        for r in reqs:
            assert r.result is not None and (r.faults == 0).all()
            d = Data(models[r.key], r.n_worlds, precision)
            d.set_state(r.qpos, r.qvel)
            step(models[r.key], d, r.horizon)
            q, v = d.state()
            d.close()
            say(f"request {r.rid} {r.key} {r.n_worlds}x{r.horizon}: "
                f"max |dq| {np.abs(q - r.result[0]).max():.3g}")
            assert np.array_equal(q, r.result[0]) and np.array_equal(v, r.result[1])
    finally:
        srv.close()
    assert srv.launches > len(reqs)      # horizons were split into chunks


@pytest.mark.gpu
def test_served_free_bodies_equal_alone():
    """Free-body pools (the golden stack under PGS, the 64-box pile under
    Newton) and a pool of trees with free bodies (a hinged box chain with
    free boxes) through the slot entry points: every served rollout equals
    the same rollout run alone, bit for bit."""
    spec = [("boxstack3", 20, 60, 0.0), ("boxpile64_newton", 6, 40, 0.0), ("boxstack3", 9, 45, 0.001),
            ("boxpile64_newton", 5, 30, 0.002), ("boxstack3", 30, 20, 0.003), ("chain4_boxes4", 12, 50, 0.0),
            ("chain4_boxes4", 7, 35, 0.001)]
    reqs = []
    for i, (key, n, h, t) in enumerate(spec):
        q, v = serve.scene_for(key).initial_state(n, 300 + i)
        reqs.append(serve.Request(key, np.asarray(q, float), np.asarray(v, float), h, t))
    pools = {"boxstack3": (serve.scene_for("boxstack3"), 40), "boxpile64_newton": (serve.scene_for("boxpile64_newton"), 8),
             "chain4_boxes4": (serve.scene_for("chain4_boxes4"), 24)}
    say(f"serve free bodies: {len(reqs)} requests, chunk 16")
    srv = serve.Server(pools, precision=abi.PRECISION_FP64, chunk=16, n_streams=2, hbm_budget=8 << 30)
    try:
        srv.run(reqs, time_limit=120.0)
        models = {k: p.model for k, p in srv.pools.items()}
        for r in reqs:
            assert r.result is not None and (r.faults == 0).all()
            d = Data(models[r.key], r.n_worlds, abi.PRECISION_FP64)
            d.set_state(r.qpos, r.qvel)
            step(models[r.key], d, r.horizon)
            q, v = d.state()
            d.close()
            say(f"request {r.rid} {r.key} {r.n_worlds}x{r.horizon}: max |dq| {np.abs(q - r.result[0]).max():.3g}")
            assert np.array_equal(q, r.result[0]) and np.array_equal(v, r.result[1])
    finally:
        srv.close()


@pytest.mark.gpu
def test_served_fp64_matches_oracle():
    from oracle import pyoracle
    reqs = _mix()[:3]
    pools = {r.key: (serve.scene_for(r.key), 64) for r in reqs}
    srv = serve.Server(pools, precision=abi.PRECISION_FP64, chunk=16, n_streams=2, hbm_budget=4 << 30)
    try:
        srv.run(reqs, time_limit=120.0)
    finally:
        srv.close()
    for r in reqs:
        k = min(r.n_worlds, 8)
        ow = pyoracle.OracleWorlds(serve.scene_for(r.key), k, 0)
        ow.set_state(r.qpos[:k], r.qvel[:k])
        ow.step(r.horizon, threads=k)
        qr, vr = ow.state()
        ow.close()
        err = np.abs(r.result[0][:k] - qr).max() / max(np.abs(qr).max(), 1.0)
        say(f"request {r.rid} {r.key}: qpos rel err vs oracle {err:.3g}")
        assert err <= 1e-9


@pytest.mark.gpu
def test_server_rejects_what_cannot_be_served():
    sc = serve.scene_for("pendulum3")
    srv = serve.Server({"pendulum3": (sc, 16)}, precision=abi.PRECISION_FP64, chunk=8, n_streams=1)
    try:
        q, v = sc.initial_state(17, 1)
        with pytest.raises(ValueError):
            srv.run([serve.Request("pendulum3", np.asarray(q), np.asarray(v), 10)])
        with pytest.raises(KeyError):
            srv.run([serve.Request("quadruped18", np.asarray(q), np.asarray(v), 10)])
        with pytest.raises(ValueError):
            srv.run([serve.Request("pendulum3", np.asarray(q)[:4], np.asarray(v)[:4], 0)])
    finally:
        srv.close()
    with pytest.raises(MemoryError):
        serve.Server({"pendulum3": (sc, 1024)}, precision=abi.PRECISION_FP64, hbm_budget=1 << 10)


@pytest.mark.gpu
def test_lazy_pools_admit_as_memory_frees():
    """Pools allocated on demand under a device budget that holds only one of
    them at a time: the quadruped's requests wait until the pendulum pool is
    idle and freed, then run; every result still equals the request stepped
    alone, and pools were created and evicted along the way."""
    reqs = _mix()
    keys = ["quadruped18", "pendulum3", "chain8"]
    specs = {k: (serve.scene_for(k), 48 if k == "quadruped18" else 64 if k == "pendulum3" else 32) for k in keys}
    probe = serve.Server(specs, precision=abi.PRECISION_MIXED, chunk=16, n_streams=2, lazy=True)
    need = {k: probe._probe(k)[1] * specs[k][1] for k in keys}
    probe.close()
    budget = max(need.values()) + min(need.values()) // 2   # one pool at a time
    say(f"lazy pools: bytes {need}, budget {budget}")
    srv = serve.Server(specs, precision=abi.PRECISION_MIXED, chunk=16, n_streams=2, hbm_budget=budget, lazy=True)
    try:
        srv.run(reqs, time_limit=120.0)
        say(f"created {srv.created}, evicted {srv.evicted}, launches {srv.launches}")
        assert srv.created >= len(keys) and srv.evicted >= 1
        assert srv.device_bytes <= budget
        for r in reqs:
            assert r.result is not None and (r.faults == 0).all()
            m = Model(serve.scene_for(r.key))
            d = Data(m, r.n_worlds, abi.PRECISION_MIXED)
            d.set_state(r.qpos, r.qvel)
            step(m, d, r.horizon)
            q, v = d.state()
            d.close()
            m.close()
            assert np.array_equal(q, r.result[0]) and np.array_equal(v, r.result[1])
    finally:
        srv.close()
