"""Serving loop on the GPU: each scene's observations are the same whether it
is served alone or continuously batched with other scenes (bit-exact: the
frame's Jacobi sums are integers and every reduction has a fixed order),
admission under a memory budget, and the real-time Poisson driver."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

PHYS = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
            environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
            volumeComplianceXPBD=0.0, neoHookean=True, globalVolumeCompliance=1.0 / 5e4,
            globalDeviatoricCompliance=1.0 / 1e4, globalKsDrag=1.0)


def _classes(pkg):
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 1500, seed=21), sc.solid("box", 3000, seed=22)]
    tips = np.stack([np.repeat((sc.top_point(m) + np.array([0.0, 0.02, 0.0]))[None], 4, axis=0) for m in meshes])
    return meshes, tips.astype(np.float32)


def _server(pkg, gpu, capacity=(4, 4), budget=1e12, max_batch=4, **kw):
    S = pkg.serve
    meshes, tips = _classes(pkg)
    kw.setdefault("max_inflight", 3)
    kw.setdefault("streams", 2)
    return S.SceneServer(meshes, capacity, budget, PHYS, tips, device=gpu, resolution=32, max_batch=max_batch,
                         grasp_radius=0.08, **kw)


def _script(pkg, actors, steps=6, seed=3):
    """Per actor: reset, `steps` steps (clamp closed on the first), and for odd
    actors a reset and two more steps."""
    S = pkg.serve
    rng = np.random.default_rng(seed)
    out = {}
    for a in actors:
        reqs = [S.Request(a, S.REQ_RESET)]
        for k in range(steps):
            d = rng.normal(0.0, 0.01, 3)
            reqs.append(S.Request(a, S.REQ_STEP, (d[0], d[1], d[2], S.CLAMP_CLOSE if k == 0 else S.CLAMP_KEEP)))
        if a % 2:
            reqs.append(S.Request(a, S.REQ_RESET))
            for k in range(2):
                reqs.append(S.Request(a, S.REQ_STEP, (0.0, 0.02, 0.0, S.CLAMP_CLOSE if k == 0 else S.CLAMP_KEEP)))
        for i, r in enumerate(reqs):
            r.seq = i
        out[a] = reqs
    return out


def _run(server, scripts, classes):
    got = {}

    def record(c, srv):
        if srv is None:
            return
        got.setdefault(c.actor, []).append((c.seq, srv.depth[c.slot].clone(), srv.obs[c.slot].clone()))

    server.on_complete = record
    for a, reqs in scripts.items():
        server.register(a, classes[a])
    # interleave the actors' requests
    longest = max(len(r) for r in scripts.values())
    for k in range(longest):
        for a, reqs in scripts.items():
            if k < len(reqs):
                server.submit(reqs[k])
    server.run_until_idle()
    torch.cuda.synchronize()
    return got


@pytest.mark.parametrize("shared", ["off", "on"])
def test_served_alone_equals_in_mix(pkg, gpu, shared):
    actors = list(range(6))
    classes = {a: a % 2 for a in actors}
    scripts = _script(pkg, actors)
    mix = _server(pkg, gpu, shared=shared)
    got_mix = _run(mix, scripts, classes)
    assert mix.stats["ticks"] < sum(len(r) for r in scripts.values())       # requests were batched
    assert mix.stats["tick_requests"] == sum(len(r) for r in scripts.values())
    for a in actors:
        alone = _server(pkg, gpu, max_batch=1, shared=shared)
        got = _run(alone, {a: scripts[a]}, classes)
        assert [s for s, _, _ in got[a]] == list(range(len(scripts[a])))
        assert [s for s, _, _ in got_mix[a]] == list(range(len(scripts[a])))
        for (s, d1, o1), (_, d2, o2) in zip(got[a], got_mix[a]):
            assert torch.equal(d1, d2), (a, s)
            assert torch.equal(o1, o2), (a, s)
        depth_last, obs_last = got[a][-1][1], got[a][-1][2]
        assert bool(torch.isfinite(obs_last).all())
        assert float((depth_last > 0).float().mean()) > 0.05                 # the scene is in view
        assert obs_last[25] > 0                                              # clamp closed: vertices grasped
        print(f"\n[serve] actor {a}: {len(got[a])} requests identical alone and in the mix; "
              f"grasped {int(obs_last[25])}, max displacement {float(obs_last[30]):.4f}")


def _observation_f64(x, v, drag, x0, tips, ground):
    """The 32 observation scalars of one scene in float64 (the layout of
    ffsrl_observe_envs)."""
    x, v, x0 = (np.asarray(a, np.float64)[:, :3] for a in (x, v, x0))
    g = drag != 0
    d = np.linalg.norm(x - x0, axis=1)
    o = np.zeros(32)
    o[0:12] = tips
    o[12:15] = x.mean(0)
    o[15:18] = x.min(0)
    o[18:21] = x.max(0)
    o[21:24] = v.mean(0)
    o[24] = np.linalg.norm(v, axis=1).max()
    o[25] = g.sum()
    o[26:29] = x[g].mean(0) if g.any() else 0.0
    o[29] = (x[:, 1] <= ground).sum()
    o[30] = d.max()
    o[31] = d.mean()
    return o


def test_observation_scalars_vs_float64(pkg, gpu):
    """Every slot's observation equals a float64 evaluation of the same state:
    counts, minima and maxima exactly (the maxima of norms to float rounding),
    sums and means within 2e-6 of the scene's scale."""
    actors = list(range(6))
    classes = {a: a % 2 for a in actors}
    srv = _server(pkg, gpu)
    _run(srv, _script(pkg, actors, seed=5), classes)
    m = srv.model
    obs = srv.obs.cpu().numpy()
    x, v, x0 = m.x.cpu().numpy(), m.v.cpu().numpy(), srv.x0.cpu().numpy()
    drag, tips = m.drag.cpu().numpy(), m.laparoscopeTip.cpu().numpy().reshape(-1, 12)
    for a in actors:
        slot = srv.actor_slot[a]
        sl = m.env_slice(slot)
        ref = _observation_f64(x[sl], v[sl], drag[sl], x0[sl], tips[slot], m.groundLevel)
        got = obs[slot].astype(np.float64)
        scale = max(1.0, float(np.abs(x[sl, :3]).max()))
        exact = [*range(12), 15, 16, 17, 18, 19, 20, 25, 29]
        assert np.array_equal(got[exact], ref[exact]), (a, got[exact] - ref[exact])
        assert np.allclose(got[[24, 30]], ref[[24, 30]], rtol=1e-6, atol=0.0), (a, got[[24, 30]], ref[[24, 30]])
        approx = [12, 13, 14, 21, 22, 23, 26, 27, 28, 31]
        assert np.allclose(got[approx], ref[approx], rtol=0.0, atol=2e-6 * scale), (a, got[approx] - ref[approx])
    assert obs[[srv.actor_slot[a] for a in actors], 25].max() > 0          # some scene holds a grasp


def test_shared_ticks_equal_batched_ticks(pkg, gpu):
    """The same request stream served with the shared-env frame in every tick
    and with one workgroup per scene: identical observations."""
    actors = list(range(6))
    classes = {a: a % 2 for a in actors}
    scripts = _script(pkg, actors, seed=11)
    runs = {}
    for shared in ("off", "on"):
        srv = _server(pkg, gpu, shared=shared)
        runs[shared] = _run(srv, scripts, classes)
        if shared == "on":
            assert int(srv.model.shared_status().abs().sum()) == 0
    for a in actors:
        assert len(runs["on"][a]) == len(scripts[a])
        for (s1, d1, o1), (s2, d2, o2) in zip(runs["off"][a], runs["on"][a]):
            assert s1 == s2 and torch.equal(d1, d2) and torch.equal(o1, o2), (a, s1)


def test_partitioned_ticks_equal_batched_ticks(pkg, gpu):
    """The same request stream served with the partitioned frame in every
    tick (500 tets per part: the scene classes in 3 and 7 parts, one tick at
    a time) and with one workgroup per scene: identical observations.  The
    partitioned frame refuses ticks on more than one stream (two frames
    could then run at once)."""
    actors = list(range(6))
    classes = {a: a % 2 for a in actors}
    scripts = _script(pkg, actors, seed=13)
    ref = _run(_server(pkg, gpu, shared="off"), scripts, classes)
    srv = _server(pkg, gpu, shared="off", parts=500, max_inflight=1, streams=1)
    assert sorted(set(srv.model.parts_per_mesh)) == [3, 7]
    got = _run(srv, scripts, classes)
    assert int(srv.model.frame_status().abs().sum()) == 0
    assert int(srv.model.p_sums.abs().sum()) == 0
    for a in actors:
        assert len(got[a]) == len(scripts[a])
        for (s1, d1, o1), (s2, d2, o2) in zip(ref[a], got[a]):
            assert s1 == s2 and torch.equal(d1, d2) and torch.equal(o1, o2), (a, s1)
    with pytest.raises(ValueError, match="one stream"):
        _server(pkg, gpu, parts=500, max_inflight=2, streams=2)


def test_reset_restores_the_initial_observation(pkg, gpu):
    S = pkg.serve
    srv = _server(pkg, gpu)
    srv.register(0, 1)
    reqs = [S.Request(0, S.REQ_RESET)] + [S.Request(0, S.REQ_STEP, (0.0, 0.05, 0.0, S.CLAMP_CLOSE))] * 3 + \
        [S.Request(0, S.REQ_RESET)]
    got = _run(srv, {0: [S.Request(r.actor, r.kind, r.action, seq=i) for i, r in enumerate(reqs)]}, {0: 1})
    first, last = got[0][0], got[0][-1]
    assert torch.equal(first[1], last[1]) and torch.equal(first[2], last[2])
    assert not torch.equal(first[1], got[0][3][1])                           # the steps changed the scene
    assert float(first[2][30]) == 0.0 and float(first[2][25]) == 0.0          # at rest, nothing grasped


def test_reset_clears_the_scene_status(pkg, gpu):
    """A reset starts a new episode: it clears the slot's status word (a
    saturated or incomplete frame of the previous episode) and no other."""
    S = pkg.serve
    srv = _server(pkg, gpu)
    _run(srv, {a: [S.Request(a, S.REQ_RESET, seq=0)] for a in (0, 1)}, {0: 0, 1: 0})
    s0, s1 = srv.actor_slot[0], srv.actor_slot[1]
    srv.model.status[s0] = pkg._lib.STATUS_SATURATED
    srv.model.status[s1] = pkg._lib.STATUS_SATURATED
    torch.cuda.synchronize()
    srv.submit(S.Request(0, S.REQ_RESET, seq=1))
    srv.run_until_idle()
    torch.cuda.synchronize()
    st = srv.model.frame_status().cpu().numpy()
    assert st[s0] == 0 and st[s1] == pkg._lib.STATUS_SATURATED


def test_admission_waits_for_memory(pkg, gpu):
    S = pkg.serve
    srv = _server(pkg, gpu, capacity=(2, 4), max_batch=8)
    budget = 2 * srv.class_bytes[1] + srv.class_bytes[1] // 2                 # room for two large scenes
    srv.budget = budget
    peak = []

    def watch(c, s):
        peak.append(srv.resident_bytes)

    actors = [10, 11, 12, 13]
    for a in actors:
        srv.register(a, 1)
    srv.on_complete = watch
    for a in actors:
        srv.submit(S.Request(a, S.REQ_RESET))
        srv.submit(S.Request(a, S.REQ_STEP, (0.0, 0.01, 0.0, S.CLAMP_CLOSE)))
        srv.submit(S.Request(a, S.CLOSE))
    assert srv.stats["admission_waits"] == 2
    srv.run_until_idle()
    torch.cuda.synchronize()
    assert srv.stats["admitted"] == 4 and srv.stats["closed"] == 4
    assert max(peak) <= budget
    done = [(c.actor, c.kind) for c in srv.completions]
    for a in actors:
        assert (a, S.REQ_RESET) in done and (a, S.REQ_STEP) in done and (a, S.CLOSE) in done
    assert srv.resident_bytes == 0


def test_first_step_of_a_new_scene_is_a_reset(pkg, gpu):
    S = pkg.serve
    srv = _server(pkg, gpu)
    srv.register(5, 0)
    srv.submit(S.Request(5, S.REQ_STEP, (0.0, 0.0, 0.0, S.CLAMP_KEEP)))
    srv.run_until_idle()
    assert srv.stats["step_as_reset"] == 1
    assert srv.completions[-1].kind == S.REQ_RESET


def test_poisson_driver(pkg, gpu):
    S = pkg.serve
    meshes, tips = _classes(pkg)
    load = S.ServingLoad(rate=4000.0, actors=32, class_p=(0.6, 0.4), seed=1)
    cls, reqs = load.generate(0.5)
    srv = S.SceneServer(meshes, (32, 32), 1e12, PHYS, tips, device=gpu, resolution=32, max_batch=64)
    for a in range(load.actors):
        srv.register(a, int(cls[a]))
    out = S.serve(srv, reqs, 0.5)
    print(f"\n[serve] {out}")
    assert out["completed_steps"] + out["completed_resets"] == len(reqs)
    assert 0.0 < out["p50_step_ms"] <= out["p99_step_ms"] < 1000.0
    assert all(c.latency > 0 for c in srv.completions)


def test_workspace_drops_the_offending_component(pkg, gpu):
    """A step whose move would take the drag point (point 3) out of the
    workspace box loses that component only, and the whole instrument moves
    by the rest (transfromCartesianActionsKernelInWorkspace,
    modelDO.py:97-154: the component is dropped, not clamped).  The same
    rule holds for ffsrl_move_tools."""
    S = pkg.serve
    meshes, tips = _classes(pkg)
    p3 = tips[0, 3].astype(np.float64)
    lo = tuple(p3 - [1.0, 0.01, 1.0])
    hi = tuple(p3 + [1.0, 0.01, 1.0])
    srv = _server(pkg, gpu, workspace=(lo, hi))
    srv.register(0, 0)
    move = (0.1, 0.05, -0.2)                           # y would leave the box: dropped
    reqs = [S.Request(0, S.REQ_RESET, seq=0), S.Request(0, S.REQ_STEP, (*move, S.CLAMP_KEEP), seq=1)]
    _run(srv, {0: reqs}, {0: 0})
    slot = srv.actor_slot[0]
    got = srv.model.laparoscopeTip[slot].cpu().numpy()
    want = tips[0] + np.float32([0.1, 0.0, -0.2])
    assert np.array_equal(got, want.astype(np.float32))
    # the direct tool move of a model follows the same rule
    m = pkg.SimModelDO([meshes[0]], [0], device=gpu, tips=tips[:1], **PHYS)
    w = torch.tensor([lo, hi], dtype=torch.float32, device=gpu)
    d = torch.tensor([move], dtype=torch.float32, device=gpu)
    pkg.lib.ffsrl_move_tools(m.batch, d.data_ptr(), w[0].data_ptr(), w[1].data_ptr(), m.stream())
    torch.cuda.synchronize()
    assert np.array_equal(m.laparoscopeTip[0].cpu().numpy(), want.astype(np.float32))
