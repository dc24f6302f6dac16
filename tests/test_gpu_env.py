"""SimEnvironment, the call pattern of the reference's demo harness
(tests/testLiverRetraction.py:19-97), on the reference's liver retraction
scene (tests/golden/liver_retraction.npz).

The same seeded actions drive three runs: the environment stepped eagerly,
the environment with its frame replayed from a HIP graph, and the facade
calls the environment stands for (resetCollisionInfo, applyActions,
applyCartesianActionsInWorkspace, stepModel with the fused refit,
renderNew).  Positions, instrument points, clamp states and images agree bit
for bit after every frame (the graph is replayed 7 times, with host reads
between replays; after a change of dispatch order the frame is captured
again).  reset(env_ids) restores the listed environments only.
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "liver_retraction.npz")
# SimModelDO arguments of testLiverRetraction.py:36-45
REF = dict(simFrameRate=30, simSubsteps=8, simConstraintsSteps=1, globalKsDrag=1.0,
           globalLaparoscopeDragLookupRadius=1.0, globalKsDistance=1.0, globalKsVolume=1.0, distance="pbd",
           volume="pbd", environmentGravity=(0.0, -9.8, 0.0), workspaceLow=(-7.0, 0.7, -10.6),
           workspaceHigh=(7.0, 9.0, 7.4))
RCM, RCM_ROT, EFFECTOR = (9.036183, 35.260103, 1.567807), (0.0, 0.0, 1.0), (0.0, 10.0, 4.0)
N, FRAMES, RES = 4, 8, 32


def _place(model, dev):
    """testLiverRetraction.py:51-58: rotation centre, effector at (0, 10, 4)."""
    model.setRotationCenter(dev(np.tile(RCM, (N, 1))), dev(np.tile(RCM_ROT, (N, 1))))
    shift = np.asarray(EFFECTOR, np.float32) - model.getLaparoscopePositionsTensor().cpu().numpy()
    model.applyCartesianActions(dev(shift))


def _actions():
    rng = np.random.default_rng(5)
    out = []
    for f in range(FRAMES):
        cart = rng.uniform(-0.1, 0.1, (N, 3)).astype(np.float32)
        joint = None
        if f in (3, 4):  # close the jaws (the clamp state switches on the second call)
            joint = np.zeros((N, 5), np.float32)
            joint[:, 4] = -0.5
        out.append((joint, cart))
    return out


def _state(model, image):
    torch.cuda.synchronize()
    return (model.vertex().cpu().numpy().copy(), model.laparoscopeTip.cpu().numpy().copy(),
            model.laparoscopeDragCutHeat.cpu().numpy().copy(), model.drag.cpu().numpy().copy(),
            image.cpu().numpy().copy())


def test_environment_eager_graph_and_facade_agree(pkg, gpu):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(gpu)
    envs = [pkg.SimEnvironment.from_usd(FIXTURE, N, device=gpu, resolution=RES, useGraph=g, **REF)
            for g in (False, True)]
    model = pkg.SimModelDO.from_usd(FIXTURE, N, device=gpu, **REF)
    bvh = pkg.SimBVH(gpu, model)
    ren = pkg.WarpRaycastRendererDO(gpu, model, cameraPos=(0.0, 50.0, 50.0), lightPos=(0.0, 35.0, 50.0),
                                    lightIntensity=0.03, resolution=RES)
    ren.lookAt((0.0, 10.0, 0.0))
    integ = pkg.SimIntegratorDO(gpu)
    for e in envs:
        _place(e.simModel, dev)
    _place(model, dev)

    def run(actions, what):
        seen = False
        for f, (joint, cart) in enumerate(actions):
            got = []
            for e in envs:
                obs = e.step(actions=None if joint is None else dev(joint), actionsCartesian=dev(cart), render=True)
                assert obs["image"].shape == (N, RES, RES, 3) and obs["tool"].shape == (N, 4, 3)
                got.append(_state(e.simModel, obs["image"]))
            model.resetCollisionInfo()
            if joint is not None:
                model.applyActions(dev(joint))
            model.applyCartesianActionsInWorkspace(dev(cart))
            integ.stepModel(model, refit=True)
            got.append(_state(model, ren.renderNew(bvh, model)))
            for g in got[:2]:
                for a, b in zip(g, got[2]):
                    assert np.array_equal(a, b), f"{what} frame {f}"
            seen |= bool(got[2][4].any())
        return seen

    assert run(_actions(), "first")  # the tissue is in view
    assert envs[1].capturedGraph is not None and envs[1].captures == 1
    # a new launch shape (table dispatch order; results unchanged) makes the
    # graph environment capture its frame again; a new constraint order is
    # written into the same arrays, so the graph reads it either way
    envs[1].simIntegrator.dispatch_sorted = False
    envs[1].simModel.set_constraint_order("local", chunk=64)
    run(_actions()[:3], "re-captured")
    assert envs[1].captures == 2

    # reset of env 1 only
    e = envs[0]
    before = _state(e.simModel, e.image)
    obs = e.reset([1], render=True)
    after = _state(e.simModel, obs["image"])
    m = e.simModel
    x0 = m.initialVertex[m.env_slice(1), :3].cpu().numpy()
    assert np.array_equal(after[0].reshape(N, -1, 3)[1], x0)
    for k in (0, 2, 3):
        assert np.array_equal(after[0].reshape(N, -1, 3)[k], before[0].reshape(N, -1, 3)[k])
    assert np.array_equal(after[1].reshape(N, -1)[1], m.initialTips[1].cpu().numpy().reshape(-1))
    assert after[2].reshape(N, 3)[1].tolist() == [0, 0, 0]
    assert not after[3].reshape(N, -1)[1].any()
    assert np.array_equal(after[4][0], before[4][0]) and np.array_equal(after[4][2:], before[4][2:])
    with pytest.raises(IndexError):
        e.reset([N])
    for env in envs:
        env.close()


def test_failed_capture_leaves_the_stream_usable(pkg, gpu):
    """A frame that fails inside the graph capture (here an invalid block
    size, rejected by the library before any launch) ends the capture and
    drops the partial graph (ffsrl_graph_abort); the error reaches the
    caller, no graph is kept, and the next step captures and replays a
    frame equal to an eager one."""
    envs = [pkg.SimEnvironment.from_usd(FIXTURE, 2, device=gpu, resolution=RES, useGraph=g, **REF)
            for g in (True, False)]
    g, eager = envs
    g.simIntegrator.block = 333
    with pytest.raises(pkg.FFSRLError, match="block must be"):
        g.step()
    assert g.capturedGraph is None and g._graph_key is None
    g.simIntegrator.block = 1024
    for _ in range(2):
        g.step()
        eager.step()
    torch.cuda.synchronize()
    assert g.captures == 1
    assert torch.equal(g.simModel.x, eager.simModel.x) and torch.equal(g.simModel.v, eager.simModel.v)
    for env in envs:
        env.close()
