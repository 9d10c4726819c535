"""Throughput of the batched newt step on one MI355X.

A step is one World::step of every world in a batch (newt/src/world.rs:
1459-1537 per world).  The headline workload is the 18-DoF floating-base
quadruped on a plane: 8192 worlds, PGS 50 iterations, Euler dt 2 ms, seeded
states dropped from 0.3-0.5 m.  Every run first steps the batch 300 steps
(untimed, on top of --warmup): the landing, which brings every quadruped
into contact.  The timed window is then the contact steady state, whatever
--warmup says.

`value` is the mixed step's world-steps/s over the --steps timed steps with
the state resident in HBM; the FP64 step (the drop-in) is timed the same way
and reported beside it.  Each launch advances every world `--chunk` steps;
HIP events recorded on the stream the kernel runs on bracket every launch,
so launch durations are the dominant kernel's durations.  The whole rollout
from the drop (settle + warmup + timed steps) is timed as well.

Reported beside the throughput:
  roofline        HBM fraction of the algorithmic state traffic (736 B per
                  world-step: qpos/qvel read and written, qacc written, FP64)
  roofline_fp64   FP64 fraction of newt's own operation count (the oracle's
                  counting scalar over 50 steps from the timed states)
  qacc_err        mixed step vs the FP64 oracle, one step from the timed
                  states; `err_100` the same per quantity after 100 steps
  cpu_baseline    the oracle's float instantiation (newt's f32 arithmetic)
                  stepping a sample of the same worlds on the host cores
  c1 .. c5        one summary per further workload: the 3-link pendulum
                  (256 worlds x 1000 steps, FP64), the 64-box pile (Newton
                  20), the uneven hinge-chain batch, and the serving mix,
                  whose request latency p50/p99 is `rollout_latency_ms`

    python bench.py                      # everything, a few minutes
    python bench.py --workload boxes     # one workload in detail
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6    # MI355X dense FP64 (vector)
QUADRUPED_SETTLE = 300     # steps from the drop to contact steady state
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_team_quadruped18_traffic.json")
METRIC = "world-steps/s mixed-prec vs FP64 GPU at qacc err<=1e-5; p99 rollout latency"


def bytes_per_world_step(nq, nv):
    """State read + state write + qacc write, FP64 (qpos, qvel, qacc)."""
    return 2 * (nq * 8 + nv * 8) + nv * 8


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def threads():
    return max(1, min(os.cpu_count() or 1, 16))


def percentile(xs, p):
    s = sorted(xs)
    k = min(len(s) - 1, max(0, int(round(p / 100.0 * (len(s) - 1)))))
    return s[k]


def time_gpu(model, worlds, precision, steps, chunk, q0, v0, settle=0, keep_state=False):
    """Steps `settle` untimed steps, then `steps` timed steps in launches of
    `chunk`; each launch is bracketed by HIP events on its stream and the
    window by device synchronizations.  Also times the whole rollout."""
    from tooling_amd import hip
    from tooling_amd.newt import Data, step

    d = Data(model, worlds, precision)
    d.set_state(q0, v0)
    stream = hip.Stream()
    sp = stream.handle
    n_launch = max(1, steps // chunk)
    n_settle = (settle + chunk - 1) // chunk if settle else 0
    ev = [(hip.Event(), hip.Event()) for _ in range(n_settle + n_launch)]
    hip.synchronize()
    for i in range(n_settle):
        ev[i][0].record(stream)
        step(model, d, min(chunk, settle - i * chunk), sp)
        ev[i][1].record(stream)
    hip.synchronize()
    warm_state = d.state() if keep_state else None
    t0 = time.perf_counter()
    for i in range(n_settle, n_settle + n_launch):
        ev[i][0].record(stream)
        step(model, d, chunk, sp)
        ev[i][1].record(stream)
    hip.synchronize()
    wall = time.perf_counter() - t0
    ms = [a.elapsed_ms(b) for a, b in ev]
    launch_ms = ms[n_settle:]
    faults = int((d.faults() != 0).sum())
    d.close()
    stream.close()
    total = n_launch * chunk
    gpu_ms = sum(launch_ms)
    rollout_steps = settle + total
    rollout_ms = sum(ms)
    return dict(steps=total, ms_per_step=gpu_ms / total, wall_s=wall, launch_ms=launch_ms, steps_per_launch=chunk,
                value=worlds * total / (gpu_ms / 1e3), faults=faults, warm_state=warm_state,
                rollout={"steps": rollout_steps, "ms": rollout_ms,
                         "value": worlds * rollout_steps / (rollout_ms / 1e3)})


def rel_err(a, r):
    """Per world: max |a - r| / max(max |r|, 1)."""
    import numpy as np
    return np.abs(a - r).max(axis=1) / np.maximum(np.abs(r).max(axis=1), 1.0)


def error_summary(err):
    import numpy as np
    return dict(max=float(err.max()), p99=float(np.percentile(err, 99)), median=float(np.median(err)),
                worlds=int(err.size), above_1e5=int((err > 1e-5).sum()))


def precision_errors(model, scene, qs, vs, n, precision, horizon=100):
    """One step's qacc from the given states, and qpos/qvel/qacc after
    `horizon` steps, against the FP64 oracle."""
    from oracle import pyoracle
    from tooling_amd.newt import Data, step

    d = Data(model, n, precision)
    d.set_state(qs[:n], vs[:n])
    ow = pyoracle.OracleWorlds(scene, n, 0)
    ow.set_state(qs[:n], vs[:n])
    step(model, d, 1)
    ow.step(1, threads=threads())
    out = {"qacc": error_summary(rel_err(d.qacc(), ow.qacc()))}
    step(model, d, horizon - 1)
    ow.step(horizon - 1, threads=threads())
    gq, gv = d.state()
    oq, ov = ow.state()
    out[f"after_{horizon}_steps"] = {
        "qpos": error_summary(rel_err(gq, oq)), "qvel": error_summary(rel_err(gv, ov)),
        "qacc": error_summary(rel_err(d.qacc(), ow.qacc()))}
    d.close()
    ow.close()
    return out


def flops_per_world_step(scene, qs, vs, n=16, steps=50):
    """newt's operation count per world-step (oracle counting scalar) over
    `steps` steps of the first n worlds from the given states."""
    from oracle import pyoracle
    ow = pyoracle.OracleWorlds(scene, n, 2)
    ow.set_state(qs[:n], vs[:n])
    ow.flops(reset=True)
    ow.step(steps, threads=1)
    c = ow.flops(reset=True)
    ow.close()
    return sum(c) / (n * steps), c


def cpu_rate(scene, qs, vs, n, steps, warmup=0):
    """The oracle's float instantiation (newt's f32 step) on the host cores,
    one world per task."""
    from oracle import pyoracle
    th = threads()
    ow = pyoracle.OracleWorlds(scene, n, 1)
    ow.set_state(qs[:n], vs[:n])
    if warmup:
        ow.step(warmup, threads=th)
    t0 = time.perf_counter()
    ow.step(steps, threads=th)
    dt = time.perf_counter() - t0
    ow.close()
    return n * steps / dt, dt, th


def measured_traffic(key):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes, when they were taken with the
    same workload key."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("key") != key:
        return None
    return t.get("bytes_per_launch")


# ---------------------------------------------------------------------------
def bench_quadruped(args):
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Model

    scene = scenes.quadruped()
    q0, v0 = scene.initial_state(args.worlds, args.seed)
    model = Model(scene)
    settle = QUADRUPED_SETTLE + args.warmup
    log(f"quadruped: {args.worlds} worlds, settle {QUADRUPED_SETTLE} + warmup {args.warmup}, steps {args.steps}, "
        f"{args.chunk} steps per launch")
    mixed = time_gpu(model, args.worlds, abi.PRECISION_MIXED, args.steps, args.chunk, q0, v0, settle, True)
    log(f"  mixed: {mixed['value']:.4g} world-steps/s ({mixed['ms_per_step']:.4f} ms/step); from the drop "
        f"{mixed['rollout']['value']:.4g}")
    fp64 = time_gpu(model, args.worlds, abi.PRECISION_FP64, args.steps, args.chunk, q0, v0, settle)
    log(f"  fp64:  {fp64['value']:.4g} world-steps/s ({fp64['ms_per_step']:.4f} ms/step); from the drop "
        f"{fp64['rollout']['value']:.4g}")
    qs, vs = mixed["warm_state"]
    # The plain FP64 step: newt's operation order, one lane per world
    # (bit-identical to the FP64 oracle), from the same settled states.
    plain = time_gpu(model, args.worlds, abi.PRECISION_FP64_EXACT, args.steps, args.chunk, qs, vs)
    log(f"  fp64 plain (newt's order, lane per world): {plain['value']:.4g} world-steps/s "
        f"({plain['ms_per_step']:.4f} ms/step)")
    n_err = min(512, args.worlds)
    err = precision_errors(model, scene, qs, vs, n_err, abi.PRECISION_MIXED)
    log(f"  mixed vs FP64 oracle: qacc max {err['qacc']['max']:.3g} over {n_err} worlds")
    launch_ms = sum(mixed["launch_ms"]) / len(mixed["launch_ms"])
    spl = mixed["steps_per_launch"]
    B = bytes_per_world_step(model.nq, model.nv)
    achieved = B * args.worlds * spl / (launch_ms / 1e3) / 1e9
    key = f"quadruped18/{args.worlds}/chunk{spl}/settle{settle}"
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(key),
                "kernel": "newt::k_team<float, 24, false>", "bytes_per_world_step": B,
                "launch_ms": launch_ms, "steps_per_launch": spl, "worlds": args.worlds, "traffic_key": key}
    fpw, counts = flops_per_world_step(scene, qs, vs)
    tflops = fpw * args.worlds * spl / (launch_ms / 1e3) / 1e12
    roofline_fp64 = {"bound": "fp64", "achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / FP64_PEAK_TFLOPS, "flop_per_world_step": fpw,
                     "flop_counts_add_mul_div_sqrt": counts,
                     "source": "oracle counting scalar, 16 worlds, 50 steps from the timed (settled) states"}
    cpu = None
    if not args.no_cpu:
        n = 64 * threads()
        rate, dt, th = cpu_rate(scene, qs, vs, n, args.cpu_steps)
        cpu = dict(value=rate, unit="world-steps/s", cores=th, kind="port",
                   sample=f"{n} worlds of the same batch from the GPU's settled states, {args.cpu_steps} steps; "
                          f"the oracle's float instantiation (newt's f32 step, byte-identical to newt's goldens) "
                          f"on {th} host threads, one world per task ({dt:.1f} s)")
        log(f"  cpu: {rate:.4g} world-steps/s on {th} threads")
    return dict(scene=scene, model=model, mixed=mixed, fp64=fp64, plain=plain, err=err, roofline=roofline,
                roofline_fp64=roofline_fp64, cpu=cpu, settle=settle)


def bench_pendulum(args, worlds=256, steps=1000):
    """C1: 256 worlds of the 3-link pendulum, 1000 FP64 steps in one launch."""
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Model

    scene = scenes.pendulum()
    q0, v0 = scene.initial_state(worlds, args.seed)
    model = Model(scene)
    out = {"workload": f"pendulum3 (C1): {worlds} worlds, 3-link planar revolute pendulum, no contacts, "
                       f"Euler dt 2 ms, {steps} steps in one launch, seed {args.seed}"}
    for name, prec in (("fp64", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED)):
        r = time_gpu(model, worlds, prec, steps, steps, q0, v0)
        out[name] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "faults": r["faults"]}
        log(f"pendulum {name}: {r['value']:.4g} world-steps/s")
    B = bytes_per_world_step(model.nq, model.nv)
    out["bytes_per_world_step"] = B
    out["hbm_frac"] = B * out["fp64"]["value"] / 1e9 / HBM_PEAK_GBS
    from oracle import pyoracle
    from tooling_amd.newt import Data, step
    d = Data(model, worlds, abi.PRECISION_FP64)
    d.set_state(q0, v0)
    step(model, d, steps)
    gq, gv = d.state()
    d.close()
    ow = pyoracle.OracleWorlds(scene, worlds, 0)
    ow.set_state(q0, v0)
    ow.step(steps, threads=threads())
    oq, ov = ow.state()
    ow.close()
    out["fp64_vs_oracle_after_1000_steps"] = {"qpos": float(rel_err(gq, oq).max()),
                                              "qvel": float(rel_err(gv, ov).max())}
    if not args.no_cpu:
        rate, dt, th = cpu_rate(scene, q0, v0, worlds, steps)
        out["cpu_baseline"] = dict(value=rate, unit="world-steps/s", cores=th, kind="port",
                                   sample=f"all {worlds} worlds x {steps} steps, oracle float ({dt:.2f} s)")
    model.close()
    return out


def bench_boxes(args, worlds=1024, settle=300, steps=20, chunk=10):
    """C3: 1024 worlds of 64 free boxes dropped on a plane (half extents ~
    U[0.05, 0.2]^3, density 1000, jittered 4x4x4 grid, random orientations),
    Newton 20, pyramidal, Euler dt 2 ms (World::add_body models;
    newt/src/world.rs:1705-1817, solver.rs:703-1407, newton.rs)."""
    import numpy as np
    from oracle import pyoracle
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Data, Model, step

    pgs_scene = scenes.box_pile()
    scene = scenes.with_options(pgs_scene, "_newton", solver=abi.SOLVER_NEWTON)
    q0, v0 = scene.initial_state(worlds, args.seed)
    model = Model(scene)
    out = {"workload": f"boxpile64 (C3): {worlds} worlds x 64 free boxes on a plane, Newton 20 pyramidal, Euler "
                       f"dt 2 ms, timed after {settle} steps, seed {args.seed}"}
    runs = {}
    for name, prec in (("fp64", abi.PRECISION_FP64), ("mixed", abi.PRECISION_MIXED)):
        r = time_gpu(model, worlds, prec, steps, chunk, q0, v0, settle, keep_state=name == "fp64")
        runs[name] = r
        out[name] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "faults": r["faults"],
                     "launch_ms": sum(r["launch_ms"]) / len(r["launch_ms"])}
        log(f"boxes {name}: {r['value']:.4g} world-steps/s ({r['ms_per_step']:.3f} ms/step)")
    qs, vs = runs["fp64"]["warm_state"]
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
    out["fp64_vs_oracle_one_step"] = {"qpos": float(rel_err(gq, oq).max()), "qvel": float(rel_err(gv, ov).max()),
                                      "qacc": float(rel_err(ga, oa).max()), "worlds": k,
                                      "contacts_min_max": [int(nc.min()), int(nc.max())]}
    fpw, counts = flops_per_world_step(scene, qs, vs, n=4, steps=5)
    launch_ms = out["fp64"]["launch_ms"]
    tflops = fpw * worlds * chunk / (launch_ms / 1e3) / 1e12
    out["roofline_fp64"] = {"achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": tflops / FP64_PEAK_TFLOPS, "flop_per_world_step": fpw,
                            "source": "oracle counting scalar, 4 worlds from the settled states, 5 steps"}
    B = bytes_per_world_step(model.nq, model.nv)
    out["hbm_frac"] = B * out["fp64"]["value"] / 1e9 / HBM_PEAK_GBS
    if not args.no_cpu:
        n = 2 * threads()
        rate, dt, th = cpu_rate(scene, qs, vs, n, 10)
        out["cpu_baseline"] = dict(value=rate, unit="world-steps/s", cores=th, kind="port",
                                   sample=f"{n} worlds from the GPU's settled states x 10 steps, oracle float "
                                          f"(newt's f32 step) ({dt:.1f} s)")
        log(f"boxes cpu: {rate:.4g} world-steps/s")
    model.close()
    return out


def bench_chains(args, worlds=16384, settle=100, steps=100):
    """C4: 16384 worlds of fixed-root hinge chains, n ~ U{2..64} links (one
    model per length), PGS 30, dt 1 ms, contacts against the plane, stepped
    as one uneven batch (newt_batch): one persistent launch per kernel
    instantiation, replayed from a HIP graph, in the launch shape the
    autotuner chose (tooling_amd/tuning/chains.json).  Timed with HIP events
    on the replay stream, the device synchronized on both sides."""
    import numpy as np
    from oracle import pyoracle
    from tooling_amd import abi, autotune, hip
    from tooling_amd.newt import Batch, Data, step

    cfg = autotune.load()
    chunk = cfg["chunk"]
    log(f"chains: {worlds} worlds, settle {settle}, steps {steps}, tuned shape {cfg}")

    def run(precision):
        lengths, datas = autotune.chain_batch(worlds, args.seed, precision)
        stream = hip.Stream()
        b = Batch([d for (_, _, _, d) in datas], order=cfg["order"], wpb_cap=tuple(cfg["wpb_cap"]),
                  use_graph=cfg["use_graph"], branches=cfg["branches"])
        b.step(settle, stream.handle)
        b.step(chunk, stream.handle)  # captures the timed graph
        hip.synchronize()
        warm = [d.state() for (_, _, _, d) in datas]
        n = max(1, steps // chunk)
        t0, t1 = hip.Event(), hip.Event()
        t0.record(stream)
        for _ in range(n):
            b.step(chunk, stream.handle)
        t1.record(stream)
        hip.synchronize()
        ms = t0.elapsed_ms(t1)
        faults = sum(int((d.faults() != 0).sum()) for (_, _, _, d) in datas)
        launches = b.launches()
        b.close()
        stream.close()
        return lengths, datas, warm, ms, n * chunk, faults, launches

    lengths, mixed, warm, ms_m, total_steps, f_mixed, launches = run(abi.PRECISION_MIXED)
    errs = []
    for (n, sc, m, d), (qs, vs) in zip(mixed, warm):
        if n % 8:
            continue
        k = min(4, len(qs))
        dd = Data(m, k, abi.PRECISION_MIXED)
        dd.set_state(qs[:k], vs[:k])
        step(m, dd, 1)
        a = dd.qacc()
        dd.close()
        ow = pyoracle.OracleWorlds(sc, k, 0)
        ow.set_state(qs[:k], vs[:k])
        ow.step(1, threads=k)
        errs.extend(rel_err(a, ow.qacc()).tolist())
        ow.close()
    for (_, _, m, d) in mixed:
        d.close()
        m.close()
    _, fp64, _, ms_f, _, f_fp64, _ = run(abi.PRECISION_FP64)
    for (_, _, m, d) in fp64:
        d.close()
        m.close()
    out = {"workload": f"chains (C4): {worlds} worlds, fixed-root hinge chains n ~ U{{2..64}} (mean "
                       f"{float(lengths.mean()):.1f}), {len(mixed)} models in one uneven batch, PGS 30, Euler "
                       f"dt 1 ms, timed after {settle} steps, seed {args.seed}",
           "launch_shape": dict(cfg, launches_per_replay=launches,
                                source="tooling_amd/tuning/chains.json (tooling_amd.autotune)")}
    out["mixed"] = {"value": worlds * total_steps / (ms_m / 1e3), "ms_per_step": ms_m / total_steps,
                    "faults": f_mixed}
    out["fp64"] = {"value": worlds * total_steps / (ms_f / 1e3), "ms_per_step": ms_f / total_steps,
                   "faults": f_fp64}
    out["speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["qacc_err_mixed"] = {"max": float(max(errs)), "worlds": len(errs)}
    B = 40.0 * float(lengths.mean())
    out["hbm_frac"] = B * out["mixed"]["value"] / 1e9 / HBM_PEAK_GBS
    log(f"chains mixed {out['mixed']['value']:.4g}, fp64 {out['fp64']['value']:.4g} world-steps/s")
    return out


def bench_serve(args, n_requests=32):
    """C5: rollout requests arriving as a Poisson process (`--rate` per
    second): model uniform over the C1-C4 families (the 3-link pendulum, the
    quadruped, the 64-box pile under Newton, fixed-root hinge chains of
    U{2..64} links), worlds ~ U{1..512}, horizon ~ U{100..2000} steps;
    continuous batching over pools allocated on demand under a 200 GB
    device budget (idle pools freed when the budget is short), mixed then
    FP64.  Latency is arrival to results on the host."""
    import numpy as np
    from tooling_amd import abi, serve

    def requests():
        return serve.rollout_mix(n_requests, args.seed, rate=args.rate,
                                 families=("pendulum3", "quadruped18", "boxes", "chain"), box_worlds=(1, 512))

    reqs = requests()
    keys = sorted({r.key for r in reqs})
    total = sum(r.n_worlds * r.horizon for r in reqs)
    log(f"serve: {len(reqs)} requests over {len(keys)} models at {args.rate}/s, {total} world-steps")
    pools = {k: (serve.scene_for(k), 2048 if k.startswith("chain") else 1024 if k.startswith("boxpile") else 8192)
             for k in keys}
    out = {"workload": f"serve (C5): {len(reqs)} rollout requests, Poisson {args.rate}/s, model ~ U{{pendulum3, "
                       f"quadruped18, boxpile64 (Newton), chain2..64}}, worlds ~ U{{1..512}}, horizon ~ "
                       f"U{{100..2000}}, pools on demand under a 200 GB budget, seed {args.seed}",
           "world_steps": total, "models": len(keys)}
    for name, precision in (("mixed", abi.PRECISION_MIXED), ("fp64", abi.PRECISION_FP64)):
        srv = serve.Server(pools, precision=precision, chunk=400, n_streams=4, quantum=0.05, lazy=True,
                           hbm_budget=200 << 30)
        try:
            srv.run([serve.Request(k, r.qpos[:1], r.qvel[:1], 1) for k in keys
                     for r in [next(x for x in reqs if x.key == k)]])
            srv.launches = srv.world_steps = 0
            rs = requests()
            elapsed = srv.run(rs, time_limit=600.0)
            faults = int(sum(int((r.faults != 0).sum()) for r in rs))
            codes = np.concatenate([r.faults[r.faults != 0] for r in rs] + [np.zeros(0, dtype=np.int32)])
            by_family = {}
            for r in rs:
                fam = "chain" if r.key.startswith("chain") else r.key
                by_family.setdefault(fam, []).append(r.latency * 1e3)
            out[name] = {"value": total / elapsed, "elapsed_s": elapsed, "latency_ms": serve.latency_summary(rs),
                         "latency_p99_ms_by_family": {f: float(np.percentile(v, 99)) for f, v in by_family.items()},
                         "launches": srv.launches, "device_bytes": srv.device_bytes, "faults": faults,
                         "fault_codes": {int(c): int((codes == c).sum()) for c in np.unique(codes)},
                         "pools_created": srv.created, "pools_evicted": srv.evicted}
        finally:
            srv.close()
        log(f"serve {name}: {out[name]['value']:.4g} world-steps/s, p50 {out[name]['latency_ms']['p50_ms']:.1f} ms, "
            f"p99 {out[name]['latency_ms']['p99_ms']:.1f} ms")
    out["speedup_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    return out


def headline_line(args, q, extras):
    mixed, fp64 = q["mixed"], q["fp64"]
    scene = q["scene"]
    serve = extras.get("c5_serve")
    lat = None
    if serve:
        l = serve["mixed"]["latency_ms"]
        lat = {"p50": l["p50_ms"], "p99": l["p99_ms"], "max": l["max_ms"], "requests": l["n"],
               "source": "c5_serve: arrival to results on the host, mixed step"}
    out = {
        "metric": METRIC,
        "value": mixed["value"],
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": mixed["steps"],
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state and factor; f32 Delassus rows, A = Y Y^T on MFMA and PGS sweeps",
        "data": "synthetic",
        "config": {"workload": f"{scene.name}: 18-DoF floating-base quadruped on a plane, {args.worlds} worlds, "
                               f"PGS {scene.options['iterations']} iterations, Euler dt {scene.options['dt']} s, "
                               f"seed {args.seed}; timed in contact steady state after {QUADRUPED_SETTLE} "
                               f"settle steps (the drop and landing) + {args.warmup} warmup steps, untimed",
                   "worlds": args.worlds, "steps_per_launch": args.chunk, "settle_steps": QUADRUPED_SETTLE,
                   "parallelism": "1 GPU, one wavefront per world"},
        "fp64": {"value": fp64["value"], "ms_per_step": fp64["ms_per_step"], "dtype": "f64",
                 "launch_ms": sum(fp64["launch_ms"]) / len(fp64["launch_ms"]), "rollout_from_drop": fp64["rollout"]},
        "fp64_plain": {"value": q["plain"]["value"], "ms_per_step": q["plain"]["ms_per_step"], "dtype": "f64",
                       "faults": q["plain"]["faults"],
                       "what": "the plain FP64 step: newt's operation order (bit-identical to the FP64 oracle), "
                               "one lane per world, timed from the same settled states"},
        "speedup_vs_fp64": mixed["value"] / fp64["value"],
        "speedup_vs_fp64_plain": mixed["value"] / q["plain"]["value"],
        "rollout_from_drop": mixed["rollout"],
        "qacc_err": q["err"]["qacc"],
        "err_per_quantity": q["err"],
        "rollout_latency_ms": lat,
        "faults": {"mixed": mixed["faults"], "fp64": fp64["faults"], "fp64_plain": q["plain"]["faults"]},
        "roofline": q["roofline"],
        "roofline_fp64": q["roofline_fp64"],
        "cpu_baseline": q["cpu"],
    }
    out.update(extras)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed steps after the intrinsic 300-step settle")
    ap.add_argument("--worlds", type=int, default=8192)
    ap.add_argument("--chunk", type=int, default=10, help="steps per kernel launch")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="the quadruped line only")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: arrivals per second")
    ap.add_argument("--workload", choices=["all", "quadruped", "pendulum", "boxes", "chains", "serve"],
                    default="all")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (the reference kernel for the HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: worlds are independent and shard by world index with no "
              "per-step exchange, so runs on more than one GPU are not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.steps % args.chunk:
        args.chunk = 1
    if args.copy_calibration:
        import torch  # a plain device copy kernel of known size
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_mib": args.copy_calibration, "copies": 5}))
        return
    from tooling_amd import hip
    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    single = {"pendulum": bench_pendulum, "boxes": bench_boxes, "chains": bench_chains, "serve": bench_serve}
    if args.workload in single:
        print(json.dumps(single[args.workload](args)))
        return
    q = bench_quadruped(args)
    extras = {}
    if args.workload == "all" and not args.no_extras:
        extras["c1_pendulum"] = bench_pendulum(args)
        extras["c3_boxes"] = bench_boxes(args)
        extras["c4_chains"] = bench_chains(args)
        extras["c5_serve"] = bench_serve(args)
    print(json.dumps(headline_line(args, q, extras)))


if __name__ == "__main__":
    main()
