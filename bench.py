"""Throughput of the batched newt step on one MI355X.

A step is one World::step of every world in the batch.  The workload is the
18-DoF floating-base quadruped on a plane: 8192 worlds, PGS 50 iterations,
dt 2 ms, Euler, seeded states, timed after a 300-step warmup that brings the
batch into contact (newt/src/world.rs:1459-1537 per world and step).

`value` is the mixed step's world-steps/s with the state resident in HBM;
the FP64 step (the drop-in) is timed in the same run and reported beside it.
Each launch advances every world `--chunk` steps; HIP events on the stream
the kernel runs on bracket every launch, so launch durations are the
dominant kernel's durations (one kernel per launch).

Reported beside the throughput:
  roofline        HBM fraction of the algorithmic state traffic (736 B per
                  world-step: qpos/qvel read and written, qacc written, FP64)
  roofline_fp64   FP64 fraction of the algorithmic FLOPs (newt's own
                  operation count, from the oracle's counting scalar on a
                  sample of the same worlds and steps)
  qacc_err        mixed step vs the FP64 oracle, one step from the timed
                  states, relative (max |Δ| / max(max |ref|, 1) per world)
  cpu_baseline    the oracle's float instantiation (newt's f32 arithmetic)
                  stepping a sample of the same worlds on the host cores

    python bench.py --gpus 1 --steps 200 --warmup 300
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_team_quadruped18_traffic.json")


def bytes_per_world_step(nq, nv):
    """State read + state write + qacc write, FP64 (qpos, qvel, qacc)."""
    return 2 * (nq * 8 + nv * 8) + nv * 8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def percentile(xs, p):
    s = sorted(xs)
    k = min(len(s) - 1, max(0, int(round(p / 100.0 * (len(s) - 1)))))
    return s[k]


def time_gpu(model, worlds, precision, steps, warmup, chunk, q0, v0, keep_state=False):
    import torch
    from tooling_amd.newt import Data, step

    d = Data(model, worlds, precision)
    d.set_state(q0, v0)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    torch.cuda.synchronize()
    if warmup:
        step(model, d, warmup, sp)
    torch.cuda.synchronize()
    warm_state = d.state() if keep_state else None
    n_launch = max(1, steps // chunk)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(n_launch)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(n_launch)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(n_launch):
        starts[i].record(stream)
        step(model, d, chunk, sp)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    faults = int(d.faults().sum())
    d.close()
    total_steps = n_launch * chunk
    gpu_ms = sum(launch_ms)
    return dict(steps=total_steps, ms_per_step=gpu_ms / total_steps, wall_s=wall, launch_ms=launch_ms,
                steps_per_launch=chunk, value=worlds * total_steps / (gpu_ms / 1e3), faults=faults,
                warm_state=warm_state)


def qacc_error(model, scene, qs, vs, n):
    """Mixed step vs the FP64 oracle, one step from the given states."""
    import numpy as np
    from oracle import pyoracle
    from tooling_amd import abi
    from tooling_amd.newt import Data, step

    d = Data(model, n, abi.PRECISION_MIXED)
    d.set_state(qs[:n], vs[:n])
    step(model, d, 1)
    a = d.qacc()
    d.close()
    ow = pyoracle.OracleWorlds(scene, n, 0)
    ow.set_state(qs[:n], vs[:n])
    ow.step(1, threads=max(1, min(os.cpu_count() or 1, 16)))
    r = ow.qacc()
    ow.close()
    err = np.abs(a - r).max(axis=1) / np.maximum(np.abs(r).max(axis=1), 1.0)
    return dict(max=float(err.max()), p99=float(np.percentile(err, 99)), median=float(np.median(err)),
                worlds=n, above_1e5=int((err > 1e-5).sum()))


def flops_per_world_step(scene, q0, v0, warmup, n=16, steps=50):
    """newt's operation count per world-step (oracle counting scalar) over
    steps warmup..warmup+steps of the first n worlds."""
    from oracle import pyoracle
    ow = pyoracle.OracleWorlds(scene, n, 2)
    ow.set_state(q0[:n], v0[:n])
    ow.step(warmup, threads=1)
    ow.flops(reset=True)
    ow.step(steps, threads=1)
    c = ow.flops(reset=True)
    ow.close()
    return sum(c) / (n * steps), c


def cpu_baseline(scene, q0, v0, warmup, steps):
    from oracle import pyoracle
    threads = max(1, min(os.cpu_count() or 1, 16))
    n = 64 * threads
    ow = pyoracle.OracleWorlds(scene, n, 1)  # float: newt's arithmetic
    ow.set_state(q0[:n], v0[:n])
    ow.step(warmup, threads=threads)
    t0 = time.perf_counter()
    ow.step(steps, threads=threads)
    dt = time.perf_counter() - t0
    ow.close()
    return dict(value=n * steps / dt, unit="world-steps/s", cores=threads, kind="port",
                sample=f"{n} worlds of the same seeded quadruped batch, steps {warmup}..{warmup + steps} after "
                       f"the same untimed warmup as the GPU; the oracle's float instantiation (newt's f32 "
                       f"step, byte-identical to newt's goldens) on {threads} host threads, one world per "
                       f"task ({dt:.1f} s)")


def measured_traffic(worlds, chunk):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes, when they were taken on this workload."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("worlds") != worlds or t.get("steps_per_launch") != chunk:
        return None
    return t.get("bytes_per_launch")


def bench_chains(args):
    """16384 worlds of fixed-root hinge chains, n ~ U{2..64} links (one model
    per length), PGS 30, dt 1 ms; every model's batch advances `chunk` steps
    per launch, launches spread over four streams.  Timed with the device
    synchronized on both sides of the K steps."""
    import numpy as np
    import torch
    from oracle import pyoracle
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Data, Model, step

    worlds = args.worlds if args.worlds != 8192 else 16384
    rng = np.random.Generator(np.random.PCG64(args.seed))
    lengths = rng.integers(2, 65, size=worlds)
    groups = [(n, int((lengths == n).sum())) for n in range(2, 65) if (lengths == n).sum() > 0]
    streams = [torch.cuda.Stream() for _ in range(4)]
    log(f"bench chains: {worlds} worlds in {len(groups)} models, warmup {args.warmup}, steps {args.steps}")

    def run(precision):
        batches = []
        for n, cnt in groups:
            sc = scenes.hinge_chain(n)
            m = Model(sc)
            d = Data(m, cnt, precision)
            q0, v0 = sc.initial_state(cnt, args.seed + n)
            d.set_state(q0, v0)
            batches.append((n, sc, m, d, q0, v0))
        torch.cuda.synchronize()
        for i, (_, _, m, d, _, _) in enumerate(batches):
            step(m, d, args.warmup, streams[i % 4].cuda_stream)
        torch.cuda.synchronize()
        warm = [d.state() for (_, _, _, d, _, _) in batches]
        t0 = time.perf_counter()
        for _ in range(args.steps // args.chunk):
            for i, (_, _, m, d, _, _) in enumerate(batches):
                step(m, d, args.chunk, streams[i % 4].cuda_stream)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        faults = sum(int(d.faults().sum()) for (_, _, _, d, _, _) in batches)
        return batches, warm, dt, faults

    mixed, warm, t_mixed, f_mixed = run(abi.PRECISION_MIXED)
    total = worlds * (args.steps // args.chunk) * args.chunk
    log(f"mixed: {total / t_mixed:.4g} world-steps/s")
    errs = []
    for (n, sc, m, d, _, _), (qs, vs) in zip(mixed, warm):
        k = min(4, len(qs))
        dd = Data(m, k, abi.PRECISION_MIXED)
        dd.set_state(qs[:k], vs[:k])
        step(m, dd, 1)
        a = dd.qacc()
        dd.close()
        ow = pyoracle.OracleWorlds(sc, k, 0)
        ow.set_state(qs[:k], vs[:k])
        ow.step(1, threads=k)
        r = ow.qacc()
        ow.close()
        errs.extend((np.abs(a - r).max(axis=1) / np.maximum(np.abs(r).max(axis=1), 1.0)).tolist())
    for b in mixed:
        b[3].close()
    fp64, _, t_fp64, f_fp64 = run(abi.PRECISION_FP64)
    for b in fp64:
        b[3].close()
    log(f"fp64: {total / t_fp64:.4g} world-steps/s")
    mean_n = float(lengths.mean())
    B = 40.0 * mean_n  # 2*(n*8 + n*8) + n*8 bytes per world-step, FP64 state
    achieved = B * total / t_mixed / 1e9
    cpu = None
    if not args.no_cpu:
        t_cpu, ws = 0.0, 0
        threads = max(1, min(os.cpu_count() or 1, 16))
        for n, cnt in groups[::4]:
            sc = scenes.hinge_chain(n)
            k = min(cnt, threads)
            q0, v0 = sc.initial_state(k, args.seed + n)
            ow = pyoracle.OracleWorlds(sc, k, 1)
            ow.set_state(q0, v0)
            ow.step(args.warmup, threads=k)
            t0 = time.perf_counter()
            ow.step(100, threads=k)
            t_cpu += time.perf_counter() - t0
            ws += k * 100
            ow.close()
        cpu = dict(value=ws / t_cpu, unit="world-steps/s", cores=threads, kind="port",
                   sample=f"up to {threads} worlds of every 4th chain length, 100 steps after the warmup, oracle "
                          f"float (newt's f32 step), one world per host thread ({t_cpu:.1f} s)")
    out = {
        "metric": "world-steps/s mixed-prec vs FP64 GPU at qacc err<=1e-5; p99 rollout latency",
        "value": total / t_mixed, "unit": "world-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_mixed / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (f32 Delassus rows)", "data": "synthetic",
        "config": {"workload": f"chains: {worlds} worlds, fixed-root hinge chains n ~ U{{2..64}} (mean "
                               f"{mean_n:.1f}), {len(groups)} models, PGS 30, Euler dt 1 ms, seed {args.seed}",
                   "worlds": worlds, "steps_per_launch": args.chunk, "streams": 4,
                   "parallelism": "1 GPU, one wavefront per world, one launch per model per chunk"},
        "fp64": {"value": total / t_fp64, "ms_per_step": t_fp64 / args.steps * 1e3, "dtype": "f64"},
        "speedup_vs_fp64": t_fp64 / t_mixed,
        "qacc_err": {"max": float(max(errs)), "median": float(np.median(errs)), "worlds": len(errs)},
        "faults": {"mixed": f_mixed, "fp64": f_fp64},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_world_step": B, "note": "wall-clock over all models' launches"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_boxes(args):
    """C3: 1024 worlds of 64 free boxes dropped on a plane (half extents ~
    U[0.05, 0.2]^3, density 1000, jittered 4x4x4 grid, random orientations),
    Newton 20, pyramidal, Euler dt 2 ms: the free-body step (World::add_body
    models; newt/src/world.rs:1705-1817, solver.rs:703-1407, newton.rs),
    FP64 with newt's operation order.  The same batch under PGS 20, a 16x
    larger PGS batch and newt's default penalty solver are reported beside
    it."""
    import numpy as np
    from oracle import pyoracle
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Data, Model, step

    worlds = args.worlds if args.worlds != 8192 else 1024
    pgs_scene = scenes.box_pile()
    scene = scenes.with_options(pgs_scene, "_newton", solver=abi.SOLVER_NEWTON)
    q0, v0 = scene.initial_state(worlds, args.seed)
    model = Model(scene)
    log(f"bench boxes: {scene.name} x {worlds} worlds, warmup {args.warmup}, steps {args.steps}, "
        f"{args.chunk} steps per launch")
    r = time_gpu(model, worlds, abi.PRECISION_FP64, args.steps, args.warmup, args.chunk, q0, v0, keep_state=True)
    log(f"newton fp64: {r['value']:.4g} world-steps/s, {r['ms_per_step']:.3f} ms/step")
    qs, vs = r["warm_state"]
    # one step from the warm states on the GPU and on the FP64 oracle: bitwise
    k = 16
    d = Data(model, k, abi.PRECISION_FP64)
    d.set_state(qs[:k], vs[:k])
    step(model, d, 1)
    gq, gv = d.state()
    ga = d.qacc()
    nc, _, _ = d.contacts()
    d.close()
    ow = pyoracle.OracleWorlds(scene, k, 0)
    ow.set_state(qs[:k], vs[:k])
    ow.step(1, threads=k)
    oq, ov = ow.state()
    oa = ow.qacc()
    ow.close()
    exact = bool(np.array_equal(gq, oq) and np.array_equal(gv, ov) and np.array_equal(ga, oa))
    log(f"one step from the warm states vs the FP64 oracle: bit-exact {exact}; contacts {nc.min()}..{nc.max()}")
    launch_ms = sum(r["launch_ms"]) / len(r["launch_ms"])
    B = bytes_per_world_step(model.nq, model.nv)
    achieved = B * worlds * r["steps_per_launch"] / (launch_ms / 1e3) / 1e9
    fpw, counts = flops_per_world_step(scene, qs, vs, 0, n=4, steps=5)
    tflops = fpw * worlds * r["steps_per_launch"] / (launch_ms / 1e3) / 1e12
    log(f"newt's operation count {fpw:.4g} flop/world-step -> {tflops:.3g} TFLOP/s")
    # PGS 20 on the same batch, and 16x the batch (one lane per world there)
    pgs_model = Model(pgs_scene)
    rp = time_gpu(pgs_model, worlds, abi.PRECISION_FP64, args.steps, args.warmup, args.chunk, q0, v0)
    log(f"pgs fp64: {rp['value']:.4g} world-steps/s, {rp['ms_per_step']:.3f} ms/step")
    big_w = 16 * worlds
    bq, bv = pgs_scene.initial_state(big_w, args.seed)
    rb = time_gpu(pgs_model, big_w, abi.PRECISION_FP64, 20, args.warmup, 10, bq, bv)
    log(f"pgs {big_w} worlds: {rb['value']:.4g} world-steps/s, {rb['ms_per_step']:.3f} ms/step")
    # newt's default solver (penalty contacts, one lane per world) on the same batch
    pen_scene = scenes.with_options(pgs_scene, "_penalty", solver=abi.SOLVER_PENALTY)
    pen_model = Model(pen_scene)
    rpen = time_gpu(pen_model, worlds, abi.PRECISION_FP64, args.steps, args.warmup, args.chunk, q0, v0,
                    keep_state=True)
    log(f"penalty fp64: {rpen['value']:.4g} world-steps/s, {rpen['ms_per_step']:.3f} ms/step")
    pq, pv = rpen["warm_state"]
    rpb = time_gpu(pen_model, big_w, abi.PRECISION_FP64, 20, args.warmup, 10, bq, bv)
    log(f"penalty {big_w} worlds: {rpb['value']:.4g} world-steps/s, {rpb['ms_per_step']:.3f} ms/step")
    cpu = None
    cpu_pgs = None
    cpu_pen = None
    if not args.no_cpu:
        threads = max(1, min(os.cpu_count() or 1, 16))
        n = 2 * threads
        ow = pyoracle.OracleWorlds(scene, n, 1)
        ow.set_state(qs[:n], vs[:n])
        t0 = time.perf_counter()
        ow.step(10, threads=threads)
        dt = time.perf_counter() - t0
        ow.close()
        cpu = dict(value=n * 10 / dt, unit="world-steps/s", cores=threads, kind="port",
                   sample=f"{n} worlds of the same batch from the GPU's warm states, 10 steps, the oracle's "
                          f"float instantiation (newt's f32 step, byte-identical to newt's solver_newton_stack "
                          f"golden) on {threads} host threads ({dt:.1f} s)")
        log(f"cpu newton: {cpu['value']:.4g} world-steps/s on {threads} threads")
        n = 4 * threads
        ow = pyoracle.OracleWorlds(pgs_scene, n, 1)
        ow.set_state(qs[:n], vs[:n])
        t0 = time.perf_counter()
        ow.step(50, threads=threads)
        dt = time.perf_counter() - t0
        ow.close()
        cpu_pgs = dict(value=n * 50 / dt, unit="world-steps/s", cores=threads, sample=f"{n} worlds x 50 steps")
        log(f"cpu pgs: {cpu_pgs['value']:.4g} world-steps/s on {threads} threads")
        n = 4 * threads
        ow = pyoracle.OracleWorlds(pen_scene, n, 1)
        ow.set_state(pq[:n], pv[:n])
        t0 = time.perf_counter()
        ow.step(100, threads=threads)
        dt = time.perf_counter() - t0
        ow.close()
        cpu_pen = dict(value=n * 100 / dt, unit="world-steps/s", cores=threads,
                       sample=f"{n} worlds x 100 steps from the GPU's warm states")
        log(f"cpu penalty: {cpu_pen['value']:.4g} world-steps/s on {threads} threads")
    out = {
        "metric": "world-steps/s (free-body step, FP64)",
        "value": r["value"], "unit": "world-steps/s", "n_gpus": 1, "steps": r["steps"], "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"boxes (C3): {worlds} worlds x 64 free boxes on a plane (2080 candidate pairs), "
                               f"Newton 20 pyramidal, Euler dt 2 ms, seed {args.seed}",
                   "worlds": worlds, "steps_per_launch": r["steps_per_launch"],
                   "parallelism": "1 GPU, one wavefront per world (k_bodies_wave)"},
        "exact_vs_fp64_oracle": {"bitwise": exact, "worlds": k, "contacts_min_max": [int(nc.min()), int(nc.max())]},
        "pgs": {"value": rp["value"], "ms_per_step": rp["ms_per_step"], "solver": "PGS 20", "faults": rp["faults"],
                "cpu_baseline": cpu_pgs},
        "penalty": {"value": rpen["value"], "ms_per_step": rpen["ms_per_step"],
                    "solver": "penalty (newt's default): spring-damper contacts, legacy box-box manifold",
                    "kernel": "k_bodies (one lane per world)", "faults": rpen["faults"], "cpu_baseline": cpu_pen,
                    "larger_batch": {"worlds": big_w, "value": rpb["value"], "ms_per_step": rpb["ms_per_step"],
                                     "faults": rpb["faults"]}},
        "pgs_larger_batch": {"worlds": big_w, "value": rb["value"], "ms_per_step": rb["ms_per_step"],
                             "steps": rb["steps"], "warmup": args.warmup,
                             "kernel": "k_bodies (one lane per world)", "faults": rb["faults"]},
        "faults": r["faults"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_bodies_wave",
                     "bytes_per_world_step": B, "launch_ms": launch_ms,
                     "steps_per_launch": r["steps_per_launch"], "worlds": worlds},
        "roofline_fp64": {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                          "frac": tflops / FP64_PEAK_TFLOPS, "flop_per_world_step": fpw,
                          "flop_counts_add_mul_div_sqrt": counts,
                          "source": "oracle counting scalar, 4 worlds from the warm states, 5 steps"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def bench_serve(args):
    """Rollout requests arriving as a Poisson process (`--rate` per second):
    model uniform over the 3-link pendulum, the quadruped and fixed-root
    hinge chains of U{2..64} links, worlds ~ U{1..512}, horizon
    ~ U{100..2000} steps.  Served by continuous batching (one resident pool
    per model, four streams) in the mixed step, then in the FP64 step; latency
    is arrival to results on the host."""
    import numpy as np
    from tooling_amd import abi, serve

    def requests():
        return serve.rollout_mix(args.requests, args.seed, rate=args.rate)

    reqs = requests()
    keys = sorted({r.key for r in reqs})
    total = sum(r.n_worlds * r.horizon for r in reqs)
    log(f"bench serve: {len(reqs)} requests over {len(keys)} models at {args.rate}/s, {total} world-steps, "
        f"chunk <= {args.serve_chunk}, quantum {args.serve_quantum_ms} ms")
    pools = {k: (serve.scene_for(k), 2048 if k.startswith("chain") else 8192) for k in keys}
    runs = {}
    for name, precision in (("mixed", abi.PRECISION_MIXED), ("fp64", abi.PRECISION_FP64)):
        srv = serve.Server(pools, precision=precision, chunk=args.serve_chunk, n_streams=4,
                           quantum=args.serve_quantum_ms / 1e3 if args.serve_quantum_ms > 0 else None)
        try:
            # Load every kernel once before the clock starts.
            srv.run([serve.Request(k, r.qpos[:1], r.qvel[:1], 1) for k in keys
                     for r in [next(x for x in reqs if x.key == k)]])
            srv.launches = srv.world_steps = 0
            rs = requests()
            last = [time.perf_counter()]

            def progress(left, t):
                if time.perf_counter() - last[0] > 20.0:
                    last[0] = time.perf_counter()
                    log(f"  {name}: {left} requests left at {t:.1f} s")

            elapsed = srv.run(rs, time_limit=600.0, progress=progress)
            faults = int(sum(int((r.faults != 0).sum()) for r in rs))
            by_family = {}
            for r in rs:
                fam = "chain" if r.key.startswith("chain") else r.key
                by_family.setdefault(fam, []).append(r.latency * 1e3)
            runs[name] = {"value": total / elapsed, "elapsed_s": elapsed, "latency_ms": serve.latency_summary(rs),
                          "latency_p99_ms_by_family": {f: float(np.percentile(v, 99)) for f, v in by_family.items()},
                          "launches": srv.launches, "device_bytes": srv.device_bytes, "faults": faults}
        finally:
            srv.close()
        log(f"{name}: {runs[name]['value']:.4g} world-steps/s over {runs[name]['elapsed_s']:.2f} s, "
            f"p50 {runs[name]['latency_ms']['p50_ms']:.1f} ms, p99 {runs[name]['latency_ms']['p99_ms']:.1f} ms")
    m, f = runs["mixed"], runs["fp64"]
    out = {
        "metric": "world-steps/s mixed-prec vs FP64 GPU at qacc err<=1e-5; p99 rollout latency",
        "value": m["value"], "unit": "world-steps/s", "n_gpus": 1, "steps": None, "warmup": None,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (f32 Delassus rows)", "data": "synthetic",
        "config": {"workload": f"serve: {len(reqs)} rollout requests, Poisson {args.rate}/s, model ~ U{{pendulum3, "
                               f"quadruped18, chain2..64}}, worlds ~ U{{1..512}}, horizon ~ U{{100..2000}}, "
                               f"seed {args.seed}",
                   "models": len(keys), "world_steps": total, "max_steps_per_launch": args.serve_chunk,
                   "launch_quantum_ms": args.serve_quantum_ms,
                   "streams": 4, "parallelism": "1 GPU, continuous batching: one resident pool per model"},
        "latency_ms": m["latency_ms"], "latency_p99_ms_by_family": m["latency_p99_ms_by_family"],
        "launches": m["launches"], "device_bytes": m["device_bytes"], "faults": {"mixed": m["faults"],
                                                                                 "fp64": f["faults"]},
        "fp64": f, "speedup_vs_fp64": m["value"] / f["value"],
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300,
                    help="untimed steps; 300 steps take the batch through landing into contact")
    ap.add_argument("--worlds", type=int, default=8192)
    ap.add_argument("--chunk", type=int, default=10, help="steps per kernel launch (one rollout)")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["quadruped", "chains", "serve", "boxes"], default="quadruped",
                    help="quadruped: 8192 quadrupeds (the headline); chains: fixed-root hinge chains of "
                         "2..64 links, one model per length, batched over streams; serve: Poisson rollout "
                         "requests served by continuous batching; boxes: 1024 worlds of 64 free boxes, Newton 20 (C3)")
    ap.add_argument("--requests", type=int, default=64, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--serve-chunk", type=int, default=400, help="serve: most steps per launch")
    ap.add_argument("--serve-quantum-ms", type=float, default=50.0,
                    help="serve: target launch length; 0 runs every launch at --serve-chunk steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (reference kernel for the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: worlds are independent and shard by world index with no "
              "per-step exchange, so runs on more than one GPU are not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.steps % args.chunk:
        args.chunk = 1
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration, "copies": 5}))
        return
    if args.workload == "chains":
        bench_chains(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "boxes":
        bench_boxes(args)
        return
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Model

    scene = scenes.quadruped()
    q0, v0 = scene.initial_state(args.worlds, args.seed)
    model = Model(scene)
    log(f"bench: {scene.name} x {args.worlds} worlds, warmup {args.warmup}, steps {args.steps}, "
        f"{args.chunk} steps per launch")
    mixed = time_gpu(model, args.worlds, abi.PRECISION_MIXED, args.steps, args.warmup, args.chunk, q0, v0,
                     keep_state=True)
    log(f"mixed: {mixed['value']:.4g} world-steps/s, {mixed['ms_per_step']:.4f} ms/step")
    fp64 = time_gpu(model, args.worlds, abi.PRECISION_FP64, args.steps, args.warmup, args.chunk, q0, v0)
    log(f"fp64: {fp64['value']:.4g} world-steps/s, {fp64['ms_per_step']:.4f} ms/step")

    qs, vs = mixed["warm_state"]
    err = qacc_error(model, scene, qs, vs, min(512, args.worlds))
    log(f"mixed qacc error vs FP64 oracle: max {err['max']:.3g}, p99 {err['p99']:.3g} over {err['worlds']} worlds")

    launch_ms = sum(mixed["launch_ms"]) / len(mixed["launch_ms"])
    steps_per_launch = mixed["steps_per_launch"]
    B = bytes_per_world_step(model.nq, model.nv)
    achieved = B * args.worlds * steps_per_launch / (launch_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.worlds, args.chunk),
                "kernel": "newt::k_team<float, 24>", "bytes_per_world_step": B,
                "launch_ms": launch_ms, "steps_per_launch": steps_per_launch, "worlds": args.worlds}
    fpw, counts = flops_per_world_step(scene, q0, v0, args.warmup)
    tflops = fpw * args.worlds * steps_per_launch / (launch_ms / 1e3) / 1e12
    roofline_fp64 = {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS, "flop_per_world_step": fpw,
                     "flop_counts_add_mul_div_sqrt": counts,
                     "source": f"oracle counting scalar, 16 worlds, steps {args.warmup}..{args.warmup + 50}"}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(scene, q0, v0, args.warmup, args.cpu_steps)
        log(f"cpu: {cpu['value']:.4g} world-steps/s on {cpu['cores']} threads")
    lat = mixed["launch_ms"]
    out = {
        "metric": "world-steps/s mixed-prec vs FP64 GPU at qacc err<=1e-5; p99 rollout latency",
        "value": mixed["value"],
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": mixed["steps"],
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (f32 Delassus rows)",
        "data": "synthetic",
        "config": {"workload": f"{scene.name}: 18-DoF floating-base quadruped on a plane, "
                               f"{args.worlds} worlds, PGS {scene.options['iterations']} iterations, "
                               f"Euler dt {scene.options['dt']} s, seed {args.seed}, timed after {args.warmup} "
                               f"steps",
                   "worlds": args.worlds, "steps_per_launch": args.chunk,
                   "precision": "mixed: FP64 state, kinematics, dynamics, factor and PGS; rows of "
                                "Y = L^-1 J^T stored in FP32",
                   "parallelism": "1 GPU, one wavefront per world, 8 worlds per workgroup"},
        "fp64": {"value": fp64["value"], "ms_per_step": fp64["ms_per_step"], "dtype": "f64",
                 "launch_ms": sum(fp64["launch_ms"]) / len(fp64["launch_ms"])},
        "speedup_vs_fp64": mixed["value"] / fp64["value"],
        "qacc_err": err,
        "rollout_latency_ms": {"steps_per_rollout": args.chunk, "worlds": args.worlds,
                               "p50": percentile(lat, 50), "p99": percentile(lat, 99), "max": max(lat),
                               "samples": len(lat)},
        "faults": {"mixed": mixed["faults"], "fp64": fp64["faults"]},
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
