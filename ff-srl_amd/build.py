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


# Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the BVH
# builder, the constraint colouring and ordering, argument checks and error
# paths run on the CPU and are checked by tests/test_host_sanitizers.py.  GPU
# code is built without sanitizers (-fno-gpu-sanitize); the library needs
# the clang ASan runtime preloaded (sanitizer_runtime()).
SAN_LIB = os.path.join(OUT_DIR, "libffsrl_host_san.so")
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
             "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=undefined", "-fno-gpu-sanitize"]


def sanitizer_runtime() -> str:
    """Path of the clang ASan runtime that SAN_LIB links against."""
    clang = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "llvm", "bin", "clang")
    out = subprocess.run([clang, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    if out.returncode != 0 or not os.path.isabs(path) or not os.path.exists(path):
        raise RuntimeError(f"clang ASan runtime not found ({clang}: {out.stdout}{out.stderr})")
    return path


def build_host_sanitized(verbose: bool = False, jobs: int = 4) -> str:
    """The library with host code under ASan + UBSan (SAN_LIB)."""
    import tempfile
    os.makedirs(OUT_DIR, exist_ok=True)
    hipcc = _hipcc()
    base = [f for f in HIPCC_FLAGS if f != "-O3"]
    with tempfile.TemporaryDirectory(prefix="ffsrl_san_") as tmpdir:
        objs, procs = [], []
        for src in SOURCES:
            obj = os.path.join(tmpdir, src.replace(".hip", ".o"))
            cmd = [hipcc, *base, *SAN_FLAGS, *PER_FILE_FLAGS.get(src, []), "-c", os.path.join(SRC, src), "-o", obj]
            objs.append(obj)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            if len(procs) >= jobs:
                _wait(procs.pop(0), verbose)
        while procs:
            _wait(procs.pop(0), verbose)
        tmp = SAN_LIB + ".tmp"
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *SAN_FLAGS[3:], "-shared-libsan", "-o", tmp, *objs]
        _wait((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)), verbose)
        os.replace(tmp, SAN_LIB)
    return SAN_LIB


def host_sanitized_lib() -> str:
    """SAN_LIB, rebuilt when a source or header is newer than it."""
    deps = [os.path.join(SRC, f) for f in os.listdir(SRC)] + \
        [os.path.join(HERE, "..", "include", f) for f in os.listdir(os.path.join(HERE, "..", "include"))]
    if not os.path.exists(SAN_LIB) or max(os.path.getmtime(d) for d in deps) > os.path.getmtime(SAN_LIB):
        build_host_sanitized()
    return SAN_LIB


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
