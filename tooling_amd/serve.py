"""Continuous-batching rollout server over the batched newt step.

newt steps one ``World`` per call (newt/src/world.rs:1459-1537) and has no
serving layer.  This module runs many rollout requests of different models,
world counts and horizons at once, with their state resident in device
memory:

* A *pool* per model: one ``newt_data`` of ``capacity`` world slots.  With
  ``lazy=False`` every pool is allocated when the server starts, inside the
  device-memory budget.  With ``lazy=True`` a pool is allocated when the
  first request for its model arrives, if its device bytes (measured once
  per model on a one-world probe, times the capacity) fit what the budget
  has left; otherwise the least recently used idle pools (no request queued
  or running) are freed first, and if that is still not enough the requests
  wait until running pools finish and become idle.  Admission therefore
  follows the memory actually free, not a start-up reservation.
* A request (model key, initial qpos/qvel of its worlds, horizon) is admitted
  when its pool has a contiguous free slot range (first fit) and is uploaded
  with ``newt_data_set_worlds_async``.  Requests that arrive while others run
  join at their pool's next chunk boundary.
* One launch advances every occupied slot of one pool by a chunk of steps
  (parked slots are skipped inside the kernel).  The chunk is the smaller of
  the pool's chunk size and the shortest remaining horizon among its
  requests, so each request leaves exactly at its horizon.
* Launches go to ``n_streams`` streams with one chunk in flight per stream.
  Pools not yet measured go first.  Otherwise stream 0 takes the idle pool
  that will finish a request soonest: the shortest remaining horizon among
  its requests times the pool's measured time per step.  This keeps short
  rollouts' latency low.  The other streams take the idle pool with the
  most remaining work, so the longest rollouts, which bound the tail
  latency, run from the start.
* Chunks are sized to about ``quantum`` seconds of the pool's step time.
  Cheap models (the pendulum) run long chunks and a 64-link chain short
  ones, so no pool holds a stream for much longer than ``quantum``.
* A finished request's worlds are copied out, then its slots are parked and
  freed.

A world's trajectory depends neither on its slot, nor on the other worlds in
the launch, nor on how its horizon is split into launches.  Each result is
therefore the result of the same request stepped alone, to the bit, in every
precision.
"""
from __future__ import annotations

import bisect
import collections
import time
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np

from . import abi, hip
from .newt import Data, Model, step
from .scenes import Scene


@dataclass(eq=False)
class Request:
    key: str                      # pool (model) key
    qpos: np.ndarray              # [n_worlds, nq]
    qvel: np.ndarray              # [n_worlds, nv]
    horizon: int                  # World::step calls per world
    arrival: float = 0.0          # seconds after the start of Server.run
    rid: int = -1
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[tuple] = None   # (qpos, qvel) after `horizon` steps
    faults: Optional[np.ndarray] = None

    @property
    def n_worlds(self):
        return self.qpos.shape[0]

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class SlotRanges:
    """First-fit allocator of contiguous slot ranges in [0, capacity)."""

    def __init__(self, capacity: int):
        self.capacity = capacity
        self.free = [(0, capacity)]   # sorted, disjoint, coalesced (first, count)

    def alloc(self, n: int) -> int:
        for i, (f, c) in enumerate(self.free):
            if c >= n:
                if c == n:
                    del self.free[i]
                else:
                    self.free[i] = (f + n, c - n)
                return f
        return -1

    def fits(self, n: int) -> bool:
        return any(c >= n for _, c in self.free)

    def release(self, first: int, n: int):
        i = bisect.bisect_left(self.free, (first, 0))
        self.free.insert(i, (first, n))
        # coalesce with the right, then the left neighbour
        if i + 1 < len(self.free) and first + n == self.free[i + 1][0]:
            self.free[i] = (first, n + self.free[i + 1][1])
            del self.free[i + 1]
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == first:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            del self.free[i]

    @property
    def n_free(self):
        return sum(c for _, c in self.free)


class Pool:
    """One model's resident batch: a newt_data of `capacity` slots."""

    def __init__(self, key: str, scene: Scene, capacity: int, precision: int):
        self.key = key
        self.scene = scene
        self.model = Model(scene)
        self.data = Data(self.model, capacity, precision)
        self.data.park_worlds(0, capacity)
        self.data.synchronize()
        self.slots = SlotRanges(capacity)
        self.queue: collections.deque = collections.deque()
        self.active: List[list] = []   # [request, first slot, remaining steps]
        self.to_park: List[tuple] = []  # released ranges not yet parked on the device
        self.busy = False
        self.step_s: Optional[float] = None  # measured seconds per step of one launch (moving average)

    def remaining_work(self) -> float:
        """Estimated seconds of stepping left for this pool's requests."""
        rem = [a[2] for a in self.active] + [r.horizon for r in self.queue]
        if not rem:
            return 0.0
        return float("inf") if self.step_s is None else max(rem) * self.step_s

    def next_finish(self) -> float:
        """Estimated seconds until this pool's next request finishes."""
        rem = [a[2] for a in self.active]
        if self.queue:
            rem.append(self.queue[0].horizon)
        if not rem:
            return float("inf")
        return 0.0 if self.step_s is None else min(rem) * self.step_s

    def close(self):
        self.data.close()
        self.model.close()


@dataclass
class _Lane:
    stream: object
    event: object
    pool: Optional[Pool] = None
    k: int = 0
    t0: float = 0.0


class Server:
    """Continuous batching of rollout requests over resident pools.

    ``pools``: key -> (scene, capacity in worlds).  ``hbm_budget``: device
    bytes the pools may allocate together; start-up fails if they need more.
    ``chunk``: steps per launch, or with ``quantum`` (seconds) the largest
    chunk; a pool's chunk is then sized to about ``quantum`` of its measured
    step time.
    """

    def __init__(self, pools: Dict[str, tuple], precision: int = abi.PRECISION_MIXED, chunk: int = 50,
                 n_streams: int = 4, hbm_budget: int = 200 << 30, quantum: Optional[float] = None,
                 lazy: bool = False):
        if hip.device_count() < 1:
            raise RuntimeError("the serving loop needs a GPU")
        if chunk < 1:
            raise ValueError("chunk must be at least 1")
        self.precision = precision
        self.chunk = chunk
        self.quantum = quantum
        self.hbm_budget = hbm_budget
        self.specs: Dict[str, tuple] = dict(pools)
        self.lazy = lazy
        self.pools: Dict[str, Pool] = {}
        self.device_bytes = 0
        self.pool_bytes: Dict[str, int] = {}
        self.pending: Dict[str, collections.deque] = {}  # requests whose pool is not allocated
        self.last_use: Dict[str, int] = {}
        self.tick = 0
        self.created = 0
        self.evicted = 0
        self._probes: Dict[str, tuple] = {}
        try:
            if not lazy:
                for key in pools:
                    self._create(key)
                    if self.device_bytes > hbm_budget:
                        raise MemoryError(f"pools need more than the {hbm_budget} B device budget "
                                          f"({self.device_bytes} B after pool {key!r})")
        except BaseException:
            self.close()
            raise
        self.lanes = [_Lane(hip.Stream(), hip.Event(timing=False)) for _ in range(n_streams)]
        self.world_steps = 0
        self.launches = 0

    def close(self):
        for p in self.pools.values():
            p.close()
        self.pools = {}
        for m, bpw in self._probes.values():
            m.close()
        self._probes = {}

    # -- pools under the device budget -------------------------------------
    def _probe(self, key: str):
        """(model, device bytes per world slot) of a model, measured once on a
        one-world data (model tables included, so the estimate is an upper
        bound of a pool's bytes per slot)."""
        if key not in self._probes:
            m = Model(self.specs[key][0])
            d = Data(m, 1, self.precision)
            bpw = d.device_bytes()
            d.close()
            self._probes[key] = (m, bpw)
        return self._probes[key]

    def _create(self, key: str):
        scene, capacity = self.specs[key]
        p = Pool(key, scene, capacity, self.precision)
        self.pools[key] = p
        self.pool_bytes[key] = p.data.device_bytes()
        self.device_bytes += self.pool_bytes[key]
        self.created += 1
        return p

    def _idle(self, p: Pool) -> bool:
        return not p.busy and not p.active and not p.queue

    def _materialize(self):
        """Allocate the pools that pending requests wait for, oldest request
        first, freeing idle pools (least recently used first) when the budget
        is short."""
        waiting = sorted((q[0].arrival, k) for k, q in self.pending.items() if q)
        for _, key in waiting:
            need = self._probe(key)[1] * self.specs[key][1]
            if self.device_bytes + need > self.hbm_budget:
                idle = sorted((self.last_use.get(k, -1), k) for k, p in self.pools.items() if self._idle(p))
                for _, k in idle:
                    if self.device_bytes + need <= self.hbm_budget:
                        break
                    self.pools.pop(k).close()
                    self.device_bytes -= self.pool_bytes.pop(k)
                    self.evicted += 1
            if self.device_bytes + need > self.hbm_budget:
                continue  # waits for running pools to finish
            p = self._create(key)
            p.queue.extend(self.pending.pop(key))

    # --------------------------------------------------------------------
    def _finish(self, lane: _Lane, now: Callable[[], float]) -> int:
        """Account a completed chunk; copy out and release finished requests.
        Returns how many requests finished."""
        p = lane.pool
        lane.pool = None
        p.busy = False
        per_step = (now() - lane.t0) / lane.k
        p.step_s = per_step if p.step_s is None else 0.7 * p.step_s + 0.3 * per_step
        done = []
        for a in p.active:
            a[2] -= lane.k
            if a[2] == 0:
                done.append(a)
        if not done:
            return 0
        faults = p.data.faults()
        for a in done:
            r, first, _ = a
            r.result = p.data.worlds(first, r.n_worlds)
            r.faults = faults[first:first + r.n_worlds].copy()
            r.finished = now()
            # Parked on the stream of the pool's next launch, before the
            # uploads that may reuse the slots.
            p.to_park.append((first, r.n_worlds))
            p.slots.release(first, r.n_worlds)
        p.active = [a for a in p.active if a[2] > 0]
        return len(done)

    def _dispatch(self, lane: _Lane, now: float) -> bool:
        short = lane is self.lanes[0]
        key = (lambda q: q.next_finish()) if short else (lambda q: -q.remaining_work())
        best = None
        for p in self.pools.values():
            if p.busy or not (p.active or (p.queue and p.slots.fits(p.queue[0].n_worlds))):
                continue
            if best is None or key(p) < key(best):
                best = p
        if best is None:
            return False
        p, s = best, lane.stream.handle
        self.tick += 1
        self.last_use[p.key] = self.tick
        for first, n in p.to_park:
            p.data.park_worlds(first, n, s)
        p.to_park.clear()
        # Admit in arrival order while the head request fits.
        while p.queue:
            first = p.slots.alloc(p.queue[0].n_worlds)
            if first < 0:
                break
            self._admit(p, p.queue.popleft(), first, s, now)
        k = self.chunk
        if self.quantum is not None and p.step_s:
            k = max(1, min(self.chunk, int(self.quantum / p.step_s)))
        k = min(k, min(a[2] for a in p.active))
        step(p.model, p.data, k, s)
        lane.event.record(lane.stream)
        lane.pool, lane.k, lane.t0 = p, k, now
        p.busy = True
        self.launches += 1
        self.world_steps += k * sum(a[0].n_worlds for a in p.active)
        return True

    def _admit(self, p: Pool, r: Request, first: int, stream, now: float):
        p.data.set_worlds_async(first, r.qpos, r.qvel, stream)
        r.admitted = now
        p.active.append([r, first, r.horizon])

    def check(self, r: Request):
        """Raises if the request can never be served."""
        if r.key not in self.specs:
            raise KeyError(f"no pool for model {r.key!r}")
        p = self.pools.get(r.key)
        model = p.model if p is not None else self._probe(r.key)[0]
        capacity = self.specs[r.key][1]
        if r.qpos.ndim != 2 or r.qvel.ndim != 2 or r.qpos.shape != (r.n_worlds, model.nq) \
                or r.qvel.shape != (r.n_worlds, model.nv):
            raise ValueError(f"request state shapes {r.qpos.shape}, {r.qvel.shape} do not match "
                             f"model {r.key!r} (nq {model.nq}, nv {model.nv})")
        if r.n_worlds < 1 or r.n_worlds > capacity:
            raise ValueError(f"request of {r.n_worlds} worlds does not fit pool {r.key!r} "
                             f"of {capacity} slots")
        if r.horizon < 1:
            raise ValueError("horizon must be at least 1")
        if self.lazy and self._probe(r.key)[1] * capacity > self.hbm_budget:
            raise MemoryError(f"pool {r.key!r} alone needs more than the {self.hbm_budget} B budget")

    def submit(self, r: Request):
        self.check(r)
        p = self.pools.get(r.key)
        if p is not None:
            p.queue.append(r)
        else:
            self.pending.setdefault(r.key, collections.deque()).append(r)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            time_limit: float = float("inf"), progress: Optional[Callable[[int, float], None]] = None) -> float:
        """Serve `requests` (each becomes visible at its `arrival`, seconds
        after the call starts) until all have finished.  Returns the elapsed
        time.  Request ids are their index in `requests`.  `progress(n_left,
        seconds)` is called whenever requests finish."""
        order = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        for i, r in enumerate(requests):
            self.check(r)
            r.rid = i
        t0 = clock()
        nxt, left = 0, len(requests)
        while left:
            now = clock() - t0
            if now > time_limit:
                raise TimeoutError(f"{left} requests unfinished after {time_limit} s")
            while nxt < len(order) and requests[order[nxt]].arrival <= now:
                self.submit(requests[order[nxt]])
                nxt += 1
            progressed = False
            stamp = lambda: clock() - t0  # noqa: E731
            if self.pending:
                self._materialize()
            for lane in self.lanes:
                if lane.pool is not None and lane.event.query():
                    n_done = self._finish(lane, stamp)
                    left -= n_done
                    progressed = True
                    if n_done and progress is not None:
                        progress(left, stamp())
            for lane in self.lanes:
                if lane.pool is None and self._dispatch(lane, stamp()):
                    progressed = True
            if not progressed:
                time.sleep(5e-5)
        return clock() - t0


def rollout_mix(n_requests: int, seed: int, rate: float = 200.0, families=("pendulum3", "quadruped18", "chain"),
                worlds=(1, 512), horizon=(100, 2000), chain_links=(2, 64), box_worlds=(1, 64)) -> List[Request]:
    """Poisson arrivals (`rate` per second) of rollout requests: model family
    uniform over `families` ("chain" draws its length uniformly from
    `chain_links`; "boxes" is the 64-box pile under Newton, with its world
    count from `box_worlds`), world count and horizon uniform over the
    inclusive ranges.  Initial states come from each scene's seeded
    sampler."""
    from . import scenes
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    cache: Dict[str, Scene] = {}
    out = []
    for i in range(n_requests):
        fam = families[int(rng.integers(len(families)))]
        key = f"chain{int(rng.integers(chain_links[0], chain_links[1] + 1))}" if fam == "chain" else fam
        if fam == "boxes":
            key = "boxpile64_newton"
        if key not in cache:
            cache[key] = scene_for(key)
        wr = box_worlds if fam == "boxes" else worlds
        n = int(rng.integers(wr[0], wr[1] + 1))
        h = int(rng.integers(horizon[0], horizon[1] + 1))
        q, v = cache[key].initial_state(n, seed * 1000003 + i)
        out.append(Request(key, np.asarray(q, dtype=np.float64), np.asarray(v, dtype=np.float64), h,
                           float(arrivals[i])))
    return out


def scene_for(key: str) -> Scene:
    """Scene of a pool key: a name in scenes.SCENES or chain<n>."""
    from . import scenes
    if key in scenes.SCENES:
        return scenes.SCENES[key]()
    if key.startswith("chain") and key[5:].isdigit():
        return scenes.hinge_chain(int(key[5:]))
    raise KeyError(f"unknown scene {key!r}")


def latency_summary(requests: List[Request]) -> dict:
    lat = np.array([r.latency for r in requests if r.latency is not None]) * 1e3
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "n": int(lat.size)}
