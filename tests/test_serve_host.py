"""Host side of the serving loop (no GPU): the seeded Poisson request stream
and the latency summary."""
import numpy as np


def test_poisson_stream(pkg):
    S = pkg.serve
    load = S.ServingLoad(rate=20000.0, actors=4096, seed=7)
    cls, reqs = load.generate(0.25)
    n = len(reqs)
    assert abs(n - 5000) < 5 * np.sqrt(5000)                       # Poisson count, mean rate * duration
    t = np.array([r.t_arrival for r in reqs])
    assert (np.diff(t) >= 0).all() and t[0] >= 0 and t[-1] <= 0.25
    seen = set()
    resets = 0
    for r in reqs:
        if r.actor not in seen:
            assert r.kind == S.REQ_RESET                            # an actor starts with a reset
            seen.add(r.actor)
        resets += r.kind == S.REQ_RESET
    later = n - len(seen)
    assert resets - len(seen) < 0.03 * max(later, 1) + 10           # ~1% resets afterwards
    freq = np.bincount(cls, minlength=4) / len(cls)
    assert np.allclose(freq, [0.4, 0.3, 0.2, 0.1], atol=0.03)
    cls2, reqs2 = S.ServingLoad(rate=20000.0, actors=4096, seed=7).generate(0.25)
    assert np.array_equal(cls, cls2) and reqs == reqs2               # seeded


def test_summary_percentiles(pkg):
    S = pkg.serve
    comps = [S.Completion(0, S.REQ_STEP, i, 0, 0.0, (i + 1) * 1e-3) for i in range(100)]
    comps.append(S.Completion(1, S.REQ_RESET, 0, 0, 0.0, 0.5))
    out = S.summarize(comps, wall=2.0, duration=2.0)
    assert out["completed_steps"] == 100 and out["completed_resets"] == 1
    assert out["env_steps_per_s"] == 50.0
    assert abs(out["p50_step_ms"] - 50.5) < 1e-9 and abs(out["p99_step_ms"] - 99.01) < 1e-9
