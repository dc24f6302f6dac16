"""GPU parity of the instrument kinematics (modelDO.py:55-289, 1331-1404,
1570-1587) against the oracle, through the C ABI.

Tolerances: transforms and capsule points within 2e-5 (absolute, metres) of
the fp32 oracle after a sequence of actions (sin/cos/asin of the device
library and of the host libm differ in the last bits); clamp states,
workspace flags and grasp selections exactly.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

KW = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
          environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
          volumeComplianceXPBD=0.0, globalLaparoscopeDragLookupRadius=0.12)


def _scene(pkg, n_env=4, tets=2500, seed=31, tilt=True):
    """Ellipsoids, each with an instrument whose drag point is 3 cm above the
    top of the body; with `tilt` the instrument is turned by up to 0.2 rad
    about its rotation centre (its drag point then misses the body)."""
    sc = pkg.scenes
    meshes, laps = [], []
    rng = np.random.default_rng(seed)
    for i in range(n_env):
        m = sc.solid("ellipsoid", tets, seed=seed + i)
        lap = pkg.Laparoscope.aimed_at(sc.top_point(m) + np.array([0.0, 0.03, 0.0]), rcm_height=0.9)
        if tilt:
            lap.rcm_rotation = (np.pi / 2 + rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), 0.0)
        lap.jaw_angle = 60.0
        meshes.append(m)
        laps.append(lap)
    return meshes, laps


def _oracle(orc, laps):
    xf = np.concatenate([lap.stored_transforms() for lap in laps])
    pts = [lap.initial_points() for lap in laps]
    return orc.OracleInstrument(xf, np.stack([p[3] for p in pts]), np.concatenate([p[0] for p in pts]),
                                np.concatenate([p[1] for p in pts]), np.float32)


def _compare(model, o, what):
    xf = model.laparoscopeXForm.cpu().numpy().reshape(-1, 4, 4)
    tips = model.laparoscopeTip.cpu().numpy().reshape(-1, 3)
    bases = model.laparoscopeBase.cpu().numpy().reshape(-1, 3)
    e = max(float(np.abs(xf - o.xf).max()), float(np.abs(tips - o.tips).max()), float(np.abs(bases - o.bases).max()))
    print(f"[instrument {what}] max |gpu - oracle| = {e:.2e}")
    assert e <= 2e-5, what
    assert model.laparoscopeDragCutHeat.cpu().numpy().tolist() == o.dch.tolist(), what
    assert model.laparoscopeDragCutHeatUpdate.cpu().numpy().tolist() == o.dchu.tolist(), what


def test_action_sequence_matches_oracle(pkg, orc, gpu):
    meshes, laps = _scene(pkg)
    n = len(meshes)
    model = pkg.SimModelDO(meshes, list(range(n)), device=gpu, laparoscopes=laps,
                           workspaceLow=(-0.5, 0.5, -0.5), workspaceHigh=(0.5, 2.5, 0.5), **KW)
    o = _oracle(orc, laps)
    assert np.array_equal(model.laparoscopeTip.cpu().numpy().reshape(-1, 3), o.tips)
    rng = np.random.default_rng(3)
    for step in range(12):
        a = np.zeros((n, 5), np.float32)
        a[:, :4] = rng.normal(0, [0.01, 0.05, 0.02, 0.02], size=(n, 4))
        a[:, 4] = rng.choice([-0.4, 0.0, 0.3], size=n)
        model.applyActions(torch.from_numpy(a).to(gpu))
        o.actions(a)
        torch.cuda.synchronize()
        _compare(model, o, f"actions {step}")
        c = rng.normal(0, 0.01, size=(n, 3)).astype(np.float32)
        c[step % n] = 0.0                                          # an env without a Cartesian action
        if step % 2:
            model.applyCartesianActions(torch.from_numpy(c).to(gpu))
            o.cartesian(c)
        else:
            model.applyCartesianActionsInWorkspace(torch.from_numpy(c).to(gpu))
            o.cartesian(c, np.array(model.workspaceLow, np.float32), np.array(model.workspaceHigh, np.float32))
            assert model.getLaparoscopeWorkspaceViolation().cpu().numpy().tolist() == o.violation.tolist()
        torch.cuda.synchronize()
        _compare(model, o, f"cartesian {step}")
    t = np.array([[0.05 * i, 2.0, 0.0] for i in range(n)], np.float32)
    r = np.array([[1.5, 0.1 * i, 0.0] for i in range(n)], np.float32)
    model.setRotationCenter(torch.from_numpy(t).to(gpu), torch.from_numpy(r).to(gpu))
    o.set_rcm(t, r)
    torch.cuda.synchronize()
    _compare(model, o, "set rotation centre")
    # a reset restores the instrument and the clamp state
    model.resetFixedModel([1])
    torch.cuda.synchronize()
    xf = model.laparoscopeXForm.cpu().numpy().reshape(n, 7, 16)
    assert np.array_equal(xf[1], laps[1].stored_transforms().reshape(7, 16))
    assert model.laparoscopeDragCutHeat.cpu().numpy()[3:6].tolist() == [0, 0, 0]


def test_workspace_violation_flags(pkg, gpu):
    meshes, laps = _scene(pkg, 2)
    d0 = np.stack([lap.initial_points()[0][3] for lap in laps])
    # updateLaparoscopeKernel puts the drag point 0.2 jaw heights nearer the rod
    # than the construction-time one
    axis = np.stack([(lap.transforms()[0] @ [0, 0, 1, 0])[:3] for lap in laps])
    d1 = d0 - 0.2 * laps[0].jaw_height * axis
    model = pkg.SimModelDO(meshes, [0, 1], device=gpu, laparoscopes=laps,
                           workspaceLow=tuple(d1.min(axis=0) - [1.0, 0.001, 1.0]),
                           workspaceHigh=tuple(d1.max(axis=0) + 1.0), **KW)
    assert np.array_equal(model.laparoscopeTip[:, 3].cpu().numpy(), d0.astype(np.float32))
    # a zero action only recomputes the points (the kernel runs for every env)
    # and leaves the flags
    model.applyCartesianActionsInWorkspace(torch.zeros((2, 3), device=gpu))
    torch.cuda.synchronize()
    assert model.getLaparoscopeWorkspaceViolation().cpu().numpy().tolist() == [0, 0]
    d = model.laparoscopeTip[:, 3].cpu().numpy()
    assert np.allclose(d, d1, atol=2e-6)
    k = int(np.argmin(d[:, 1]))
    a = np.array([[0.0, 0.0, 0.02]] * 2, np.float32)
    a[k] = [0.01, -0.01, 0.0]
    model.applyCartesianActionsInWorkspace(torch.from_numpy(a).to(gpu))
    torch.cuda.synchronize()
    assert model.getLaparoscopeWorkspaceViolation().cpu().numpy().tolist() == [int(e == k) for e in range(2)]
    moved = model.laparoscopeTip[:, 3].cpu().numpy() - d
    assert np.allclose(moved[k], [0.01, 0.0, 0.0], atol=3e-5)
    assert np.allclose(moved[1 - k], [0.0, 0.0, 0.02], atol=3e-5)


def test_closing_the_jaws_grasps_and_drags(pkg, gpu):
    """Close (twice: the second call switches the grasp on), check the grasp
    set against |p - drag point| < lookup radius, then lift the instrument:
    the grasped vertices follow the jaw."""
    meshes, laps = _scene(pkg, 3, tets=5000, tilt=False)
    n = len(meshes)
    kw = dict(KW, globalLaparoscopeDragLookupRadius=0.2)
    model = pkg.SimModelDO(meshes, list(range(n)), device=gpu, laparoscopes=laps, **kw)
    integ = pkg.SimIntegratorDO(gpu)
    integ.stepModel(model)
    close = torch.zeros((n, 5), device=gpu)
    close[:, 4] = -0.5
    model.applyActions(close)
    assert int(model.drag.sum().item()) == 0                       # first close: still open at 60 degrees
    model.applyActions(close)
    torch.cuda.synchronize()
    assert model.laparoscopeDragCutHeat.cpu().numpy()[0::3].tolist() == [1] * n
    p = model.p[:, :3].cpu().numpy()
    drag = model.drag.cpu().numpy()
    tips = model.laparoscopeTip.cpu().numpy()
    for e in range(n):
        s = model.env_slice(e)
        want = (np.linalg.norm(p[s] - tips[e, 3], axis=1) < 0.2).astype(np.float32)
        assert want.sum() >= 5
        assert np.array_equal(drag[s], want)
    grasped = [np.nonzero(drag[model.env_slice(e)])[0] for e in range(n)]
    y0 = [model.vertex()[model.env_slice(e)][torch.as_tensor(g, device=gpu)][:, 1].mean().item()
          for e, g in enumerate(grasped)]
    lift = torch.tensor([[0.0, 0.01, 0.0]] * n, device=gpu)
    for _ in range(10):
        model.applyCartesianActions(lift)
        integ.stepModel(model)
    torch.cuda.synchronize()
    for e, g in enumerate(grasped):
        y1 = model.vertex()[model.env_slice(e)][torch.as_tensor(g, device=gpu)][:, 1].mean().item()
        assert y1 > y0[e] + 0.03
    # opening releases the grasp on the next jaw action
    opn = torch.zeros((n, 5), device=gpu)
    opn[:, 4] = 0.3
    model.applyActions(opn)
    model.applyActions(opn)
    torch.cuda.synchronize()
    assert model.laparoscopeDragCutHeat.cpu().numpy()[0::3].tolist() == [0] * n
    assert int(model.drag.sum().item()) == 0
