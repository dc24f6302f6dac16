"""bench.py's command-line contract on the host: a request for more than one
GPU is refused with one JSON line and exit status 2 before the GPU runtime
is touched, and the default run is one GPU."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    for n in (2, 8):
        r = _bench("--gpus", str(n), "--steps", "1", "--warmup", "0")
        assert r.returncode == 2, r.stderr
        lines = [l for l in r.stdout.splitlines() if l.strip()]
        assert len(lines) == 1
        out = json.loads(lines[0])
        assert out["n_gpus"] == n
        assert "one GPU" in out["error"]


def test_defaults_are_one_gpu_and_a_short_run():
    sys.path.insert(0, ROOT)
    try:
        import bench
        old = sys.argv
        sys.argv = ["bench.py"]
        try:
            a = bench.parse()
        finally:
            sys.argv = old
    finally:
        sys.path.remove(ROOT)
    assert a.gpus == 1 and a.workload == "mixed"
    assert 1 <= a.steps <= 100 and 0 <= a.warmup <= 10
    # serving defaults: partitioned ticks (1,280 tets per part), one tick in flight
    assert a.serve_inflight == 1 and a.serve_streams == 1 and a.serve_parts == 1280
