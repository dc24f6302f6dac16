"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer.

The library's host C++ (the SAH BVH builder, the constraint colouring and
ordering, argument checks and the error-message path) and the oracle's C run
tests/_host_san_driver.py twice: on the normal builds, and on builds with
ASan + UBSan (clang's runtime for the library, whose GPU code is built
without sanitizers; gcc's for the oracle), with the sanitizer runtime
preloaded into the Python process.  A sanitizer report fails the run (UBSan
is built without recovery); the output digests of both builds must match.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "_host_san_driver.py")
ENV_BASE = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
            "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def _run(what, env_extra):
    env = dict(os.environ, **env_extra)
    p = subprocess.run([sys.executable, DRIVER, what], env=env, capture_output=True, text=True, timeout=600)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    digest = [l for l in p.stdout.splitlines() if l.startswith("digest ")]
    assert len(digest) == 1, out[-2000:]
    return digest[0]


@pytest.fixture(scope="module")
def build_mod():
    import importlib
    return importlib.import_module("ff-srl_amd.build")


def test_library_host_code(build_mod):
    if not os.path.exists(build_mod.LIB):
        build_mod.build()
    build_mod.host_sanitized_lib()
    plain = _run("product", {"FFSRL_LIB": build_mod.LIB})
    rt = build_mod.sanitizer_runtime()
    san = _run("product", dict(ENV_BASE, FFSRL_LIB=build_mod.SAN_LIB, LD_PRELOAD=rt,
                               LD_LIBRARY_PATH=os.path.dirname(rt)))
    assert san == plain


def test_a_planted_overflow_is_reported(build_mod):
    """The sanitized run does catch a host-side mistake: a BVH output buffer
    one element short (tests/_host_san_driver.py short_buffer)."""
    build_mod.host_sanitized_lib()
    rt = build_mod.sanitizer_runtime()
    env = dict(os.environ, **ENV_BASE, FFSRL_LIB=build_mod.SAN_LIB, LD_PRELOAD=rt,
               LD_LIBRARY_PATH=os.path.dirname(rt))
    p = subprocess.run([sys.executable, DRIVER, "short_buffer"], env=env, capture_output=True, text=True,
                       timeout=600)
    assert p.returncode != 0 and "heap-buffer-overflow" in p.stderr, p.stderr[-3000:]


def test_oracle(orc):
    oracle_dir = os.path.dirname(orc.LIB_PATH.rstrip(os.sep))
    oracle_dir = os.path.dirname(oracle_dir)                   # oracle/_build -> oracle
    subprocess.run(["make", "-s", "-C", oracle_dir, "san"], check=True)
    san_lib = os.path.join(oracle_dir, "_build", "liboracle_san.so")
    rt = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(rt):
        pytest.skip("gcc's ASan runtime is not installed")
    plain = _run("oracle", {"FFSRL_ORACLE_LIB": orc.LIB_PATH})
    san = _run("oracle", dict(ENV_BASE, FFSRL_ORACLE_LIB=san_lib, LD_PRELOAD=rt))
    assert san == plain
