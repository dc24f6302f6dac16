"""Builds the gfx950 C-ABI library ``_build/libffsrl.so`` with hipcc.

All HIP code is built with ``-ffp-contract=off`` (no fused multiply-add, so the
refit and ray-cast arithmetic rounds exactly like a CPU evaluation) and with
packed FP32 instructions disabled.  Division and square root keep their IEEE
(correctly rounded) HIP defaults; nothing is built with fast-math.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libffsrl.so")
SOURCES = ["physics.hip", "bvh.hip", "render.hip", "meshprep.hip", "serve.hip", "instrument.hip", "bvh_build.hip"]
ARCH = os.environ.get("FFSRL_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


# The physics kernels' results are compared within a positional tolerance, so
# they may fuse multiply-adds.  -ffp-contract=on fuses only a multiply and an
# add of one source expression, so every instantiation of a kernel (block
# sizes, loop forms) rounds alike; -ffp-contract=fast fused differently in the
# 1024-thread and the 256/512-thread loops.  The refit and the ray caster
# (bit-exact bounds and hits) keep -ffp-contract=off.
PER_FILE_FLAGS = {"physics.hip": ["-ffp-contract=on"]}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build(verbose: bool = False, jobs: int = 3) -> str:
    """Compile every HIP source for gfx950 and link the shared library."""
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = _hipcc()
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(OUT_DIR, src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPCC_FLAGS, *PER_FILE_FLAGS.get(src, []), "-c", os.path.join(SRC, src), "-o", obj]
        objs.append(obj)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _wait(procs.pop(0), verbose)
    while procs:
        _wait(procs.pop(0), verbose)
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    _wait((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)), verbose)
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    return LIB


def _wait(item, verbose):
    cmd, proc = item
    out, _ = proc.communicate()
    out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose=True))
