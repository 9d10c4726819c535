"""Throughput of the batched newt step on one MI355X.

A step is one World::step of every world in the batch.  The workload is the
18-DoF floating-base quadruped on a plane (8192 worlds, PGS 50 iterations,
dt 2 ms, Euler).  `value` is the mixed-precision step's world-steps/s with
the state resident in HBM; the FP64 step (the drop-in) is timed in the same
run and reported beside it, with the ratio.  Timing uses HIP events on the
stream the kernels run on.  The CPU baseline is the oracle's float
instantiation (newt's own f32 arithmetic) stepping a bounded sample of the
same worlds on the host's cores.

    python bench.py --gpus 1 --steps 200 --warmup 20
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic HBM bytes per world-step: state read + state write + qacc write
# (qpos FP64; qvel/qacc FP64 in the FP64 step, FP32 in the mixed step).
def bytes_per_world_step(nq, nv, precision):
    if precision == 0:
        return 2 * (nq * 8 + nv * 8) + nv * 8
    return 2 * (nq * 8 + nv * 4) + nv * 4


HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_gpu(model, scene, worlds, precision, steps, warmup, chunk, q0, v0):
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
    return dict(steps=total_steps, ms_per_step=gpu_ms / total_steps, wall_s=wall,
                launch_ms=sum(launch_ms) / len(launch_ms), steps_per_launch=chunk,
                value=worlds * total_steps / (gpu_ms / 1e3), faults=faults)


def cpu_baseline(scene, q0, v0, warmup, steps=100):
    from oracle import pyoracle
    threads = max(1, min(os.cpu_count() or 1, 16))
    n = 32 * threads
    ow = pyoracle.OracleWorlds(scene, n, 1)  # float: newt's arithmetic
    ow.set_state(q0[:n], v0[:n])
    ow.step(warmup, threads=threads)
    t0 = time.perf_counter()
    ow.step(steps, threads=threads)
    dt = time.perf_counter() - t0
    ow.close()
    return dict(value=n * steps / dt, unit="world-steps/s", cores=threads, kind="port",
                sample=f"{n} worlds of the same seeded quadruped batch, steps {warmup}..{warmup + steps} "
                       f"(after the same untimed warmup as the GPU), C++ float restatement of newt's "
                       f"step on {threads} host threads ({dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300,
                    help="untimed steps; 300 steps take the batch through landing into contact")
    ap.add_argument("--worlds", type=int, default=8192)
    ap.add_argument("--chunk", type=int, default=10, help="steps per kernel launch")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: worlds are independent and shard by world index with no "
              "per-step exchange, so multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.steps % args.chunk:
        args.chunk = 1
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    from tooling_amd import abi, scenes
    from tooling_amd.newt import Model

    scene = scenes.quadruped()
    q0, v0 = scene.initial_state(args.worlds, args.seed)
    model = Model(scene)
    log(f"bench: {scene.name} x {args.worlds} worlds, warmup {args.warmup}, steps {args.steps}")
    mixed = time_gpu(model, scene, args.worlds, abi.PRECISION_MIXED, args.steps, args.warmup, args.chunk, q0, v0)
    log(f"mixed: {mixed['value']:.4g} world-steps/s, {mixed['ms_per_step']:.4f} ms/step")
    fp64 = time_gpu(model, scene, args.worlds, abi.PRECISION_FP64, args.steps, args.warmup, args.chunk, q0, v0)
    log(f"fp64: {fp64['value']:.4g} world-steps/s, {fp64['ms_per_step']:.4f} ms/step")

    B = bytes_per_world_step(model.nq, model.nv, abi.PRECISION_MIXED)
    achieved = B * args.worlds * mixed["steps_per_launch"] / (mixed["launch_ms"] / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_step<float>", "bytes_per_world_step": B,
                "launch_ms": mixed["launch_ms"], "steps_per_launch": mixed["steps_per_launch"]}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(scene, q0, v0, args.warmup)
        log(f"cpu: {cpu['value']:.4g} world-steps/s on {cpu['cores']} threads")
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
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{scene.name}: 18-DoF floating-base quadruped on a plane, "
                               f"{args.worlds} worlds, PGS {scene.options['iterations']} iterations, "
                               f"Euler dt {scene.options['dt']} s, seed {args.seed}",
                   "worlds": args.worlds, "precision": "mixed (FP32 bulk, FP64 Cholesky and qpos)",
                   "parallelism": "1 GPU, one lane per world"},
        "fp64": {"value": fp64["value"], "ms_per_step": fp64["ms_per_step"], "dtype": "f64"},
        "speedup_vs_fp64": mixed["value"] / fp64["value"],
        "faults": {"mixed": mixed["faults"], "fp64": fp64["faults"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
