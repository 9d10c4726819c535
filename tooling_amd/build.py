"""Build the HIP library (tooling_amd/_build/libnewt_gpu.so) for gfx950.

Run as ``python -m tooling_amd.build``.  hipcc cross-compiles on a host with
no GPU.  Packed FP32 instructions are disabled everywhere.  newt_gpu.hip holds
the bit-exact lane-per-world step, compiled with FP contraction off so it keeps
newt's rounding; the newt_team*.hip units hold the instantiations of the
tolerance-bound persistent step, compiled with contraction on (fused
multiply-adds), one or two per unit so they compile in parallel.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libnewt_gpu.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # falling off the end of a value-returning function is undefined
    # behaviour that the compiler may turn into a jump anywhere
    "-Werror=return-type",
]

# Per translation unit: FP contraction.
SOURCES = {"newt_gpu.hip": ["-ffp-contract=off"], "newt_noalloc.hip": ["-ffp-contract=off"],
           "newt_batch.hip": ["-ffp-contract=off"], "newt_model_json.hip": ["-ffp-contract=off"]}
for _unit in ("newt_team.hip", "newt_team24x.hip", "newt_team32.hip", "newt_team64f.hip", "newt_team64fx.hip",
              "newt_team64d.hip", "newt_team64dx.hip"):
    SOURCES[_unit] = ["-ffp-contract=fast"]


def _includes(path, seen=None):
    """Every file `path` pulls in through quoted #include lines, recursively."""
    seen = set() if seen is None else seen
    base = os.path.dirname(path)
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return seen
    for line in lines:
        line = line.strip()
        if line.startswith("#include \""):
            dep = os.path.normpath(os.path.join(base, line.split('"')[1]))
            if dep not in seen and os.path.exists(dep):
                seen.add(dep)
                _includes(dep, seen)
    return seen


def _stale(src_path, obj):
    """The object is missing or older than its source or any header it includes."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(f) > t for f in [src_path, *_includes(src_path)])


def build(verbose: bool = False, only=None) -> str:
    """Compile each translation unit (in parallel) and link the library.
    A unit is recompiled when its object is older than the unit or any header
    it includes; ``only`` names units to recompile in any case (``only`` given:
    the other up-to-date objects are reused; not given: every unit is
    recompiled)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = []
    for src, extra in SOURCES.items():
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        if only and src not in only and not _stale(os.path.join(CSRC, src), obj):
            jobs.append((src, obj, None))
            continue
        cmd = [hipcc, *compile_flags, *extra, "-c", "-o", obj, os.path.join(CSRC, src)]
        jobs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    out = []
    failed = False
    for src, obj, p in jobs:
        if p is None:
            continue
        so, se = p.communicate()
        out.append(so + se)
        if p.returncode != 0:
            sys.stderr.write(so + se)
            failed = True
    if failed:
        raise RuntimeError("hipcc failed building libnewt_gpu.so")
    tmp = LIB + ".tmp"
    res = subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _, o, _ in jobs],
                         capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking libnewt_gpu.so failed")
    if verbose:
        noise = "packed-fp32-ops"
        msg = "\n".join(l for l in "".join(out).splitlines() if noise not in l)
        if msg.strip():
            print(msg)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, only=sys.argv[1:] or None))
