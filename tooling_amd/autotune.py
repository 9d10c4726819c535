"""Autotuner of the uneven-batch step (newt_batch): picks, by measurement on
the GPU, the launch shape of a batch of trees of different sizes.

Swept: the job order (deepest trees first or as given), the most worlds per
workgroup at each register width (24, 32, 64), parallel graph branches per
kernel group, and the steps per graph replay.  Each configuration steps the
same settled batch for the same number of steps, timed with HIP events on
the stream of the replay; the fastest is written, with the whole sweep, to a
JSON table that bench.py (and any caller) reads with ``load()``.

    python -m tooling_amd.autotune            # the C4 chain batch, writes tuning/chains.json
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "chains.json")

DEFAULT = {"order": 1, "wpb_cap": [0, 0, 0], "use_graph": 1, "branches": 1, "chunk": 20}


def load(path: str = TABLE) -> dict:
    """The tuned launch shape ({order, wpb_cap, use_graph, branches, chunk});
    the defaults when no table has been written."""
    try:
        with open(path) as f:
            return dict(DEFAULT, **json.load(f)["best"])
    except (OSError, ValueError, KeyError):
        return dict(DEFAULT)


def chain_batch(worlds: int, seed: int, precision: int):
    """The uneven chain batch: n ~ U{2..64} links per world, one Data per
    length, each started from its scene's seeded sampler."""
    from . import scenes
    from .newt import Data, Model
    rng = np.random.Generator(np.random.PCG64(seed))
    lengths = rng.integers(2, 65, size=worlds)
    out = []
    for n in range(2, 65):
        cnt = int((lengths == n).sum())
        if not cnt:
            continue
        sc = scenes.hinge_chain(n)
        m = Model(sc)
        d = Data(m, cnt, precision)
        q0, v0 = sc.initial_state(cnt, seed + n)
        d.set_state(q0, v0)
        out.append((n, sc, m, d))
    return lengths, out


def measure(datas, cfg, steps, stream, log=None):
    from . import hip
    from .newt import Batch
    b = Batch([d for (_, _, _, d) in datas], order=cfg["order"], wpb_cap=tuple(cfg["wpb_cap"]),
              use_graph=cfg["use_graph"], branches=cfg["branches"])
    chunk = cfg["chunk"]
    b.step(chunk, stream.handle)  # capture and first replay, untimed
    hip.synchronize()
    t0, t1 = hip.Event(), hip.Event()
    n = max(1, steps // chunk)
    t0.record(stream)
    for _ in range(n):
        b.step(chunk, stream.handle)
    t1.record(stream)
    t1.synchronize()
    ms = t0.elapsed_ms(t1)
    launches = b.launches()
    b.close()
    return ms / (n * chunk), launches


def main():
    from . import abi, hip
    ap = argparse.ArgumentParser()
    ap.add_argument("--worlds", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--settle", type=int, default=100)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--out", default=TABLE)
    args = ap.parse_args()
    lengths, datas = chain_batch(args.worlds, args.seed, abi.PRECISION_MIXED)
    stream = hip.Stream()
    from .newt import Batch
    b = Batch([d for (_, _, _, d) in datas])
    b.step(args.settle, stream.handle)
    hip.synchronize()
    b.close()
    space = {"order": [1, 0], "wpb_cap": [[0, 0, 0], [4, 2, 2], [2, 1, 1]], "use_graph": [1],
             "branches": [1, 0], "chunk": [10, 50]}
    rows = []
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        cfg = dict(zip(keys, vals))
        ms, launches = measure(datas, cfg, args.steps, stream)
        rate = args.worlds / (ms / 1e3)
        rows.append(dict(cfg, ms_per_step=ms, world_steps_per_s=rate, launches=launches))
        print(f"{cfg}: {ms:.3f} ms/step, {rate:.4g} world-steps/s", file=sys.stderr, flush=True)
    best = min(rows, key=lambda r: r["ms_per_step"])
    table = {"workload": f"{args.worlds} fixed-root hinge chains, n ~ U{{2..64}} (seed {args.seed}), mixed "
                         f"precision, settled {args.settle} steps, timed {args.steps} steps per configuration",
             "best": {k: best[k] for k in keys}, "best_ms_per_step": best["ms_per_step"],
             "best_world_steps_per_s": best["world_steps_per_s"], "sweep": rows,
             "measured": time.strftime("%Y-%m-%d")}
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    for (_, _, m, d) in datas:
        d.close()
        m.close()
    print(json.dumps({"best": table["best"], "ms_per_step": best["ms_per_step"]}))


if __name__ == "__main__":
    main()
