"""Build libbfamd.so (the HIP solver behind include/bf/bf.h) for gfx950, and
libbfprobe.so (the benchmark's bandwidth probes, csrc/bf_probe.h).

The step kernel source is compiled once per velocity set (-DBF_VS=<id>) so
that the four sets build in parallel; the host runtime is one more object.
Floating-point contraction is off so the FP32 path keeps the operation order
of the CPU restatement; packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libbfamd.so")
PROBE_LIB = os.path.join(HERE, "libbfprobe.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
VELOCITY_SETS = {"d2q9": 2, "d3q15": 3, "d3q19": 4, "d3q27": 5}

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose=False, jobs=None, out=None, defines=()):
    """Compile every object and link libbfamd.so; returns the library path.
    `out` and `defines` (e.g. ["-DBF_NT_STORE=1"]) build a variant elsewhere."""
    lib = out or LIB
    build_dir = BUILD if out is None else out + ".objs"
    extra = list(defines)
    os.makedirs(build_dir, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    units = []
    for name, vs in VELOCITY_SETS.items():
        obj = os.path.join(build_dir, f"bf_step_{name}.o")
        units.append([HIPCC, *COMMON, *extra, f"-DBF_VS={vs}", "-c", os.path.join(CSRC, "bf_step.hip"), "-o", obj])
    for src in ("bf_solver", "bf_channel", "bf_mc", "bf_voxel", "bf_probe"):
        units.append([HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src + ".hip"),
                      "-o", os.path.join(build_dir, src + ".o")])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, units):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [u[-1] for u in units if not u[-1].endswith("bf_probe.o")]
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    if out is None:
        # the benchmark's bandwidth probes: a library of their own
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", PROBE_LIB + ".tmp",
              os.path.join(build_dir, "bf_probe.o")])
        os.replace(PROBE_LIB + ".tmp", PROBE_LIB)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
