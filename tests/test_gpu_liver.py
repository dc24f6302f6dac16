"""The reference's liver retraction scene end to end on the GPU: the scene
(tests/golden/liver_retraction.npz, converted from
FF_SRL/FF_SRL/scenes/liverRetraction.usd), the set-up of
tests/testLiverRetraction.py:44-58 (rotation centre, effector placement),
then an approach, a jaw close (grasp), a lift and a workspace-limited move,
all envs stepping together.

Parity: instrument points within 5e-5 (coordinates up to ~35); grasp
selections, clamp states and workspace flags exactly; all envs bit-identical
(they get the same actions); vertex positions within 1e-5 x the mesh
diagonal of the fp32 oracle over the horizon where the scheme is
reproducible: the set-up and the first 3 approach frames, over which the
fp32 and fp64 oracles agree to 5e-6 x diag.  From the 5th approach frame on,
the hanging free edge of this scene separates the two oracles by 1.2e-5 to
3.3e-5 x diag (8e-6 after the lift), so no float32 implementation can be
held to 1e-5 there; the positional errors are printed, not asserted.
"""
HORIZON = 3   # approach frames with positional parity asserted
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "liver_retraction.npz")
# SimModelDO arguments of testLiverRetraction.py:36-45 (the DO step runs the PBD forms, Q5)
REF = dict(simFrameRate=30, simSubsteps=8, simConstraintsSteps=1, globalKsDrag=1.0,
           globalLaparoscopeDragLookupRadius=1.0, globalKsDistance=1.0, globalKsVolume=1.0, distance="pbd",
           volume="pbd", environmentGravity=(0.0, -9.8, 0.0))
ORACLE = dict(substeps=8, iterations=1, fps=30, gravity=(0.0, -9.8, 0.0), damping=0.1, dist_mode=1, dist_k=1.0,
              vol_mode=1, vol_k=1.0, drag_ks=1.0)
RCM, RCM_ROT, EFFECTOR = (9.036183, 35.260103, 1.567807), (0.0, 0.0, 1.0), (0.0, 10.0, 4.0)
WS_LO, WS_HI = (-7.0, 0.7, -10.6), (7.0, 9.0, 7.4)


def test_liver_retraction(pkg, orc, gpu):
    n = 4
    model = pkg.SimModelDO.from_usd(FIXTURE, n, device=gpu, workspaceLow=WS_LO, workspaceHigh=WS_HI, **REF)
    mesh, lap = model.scene.meshes[0], model.scene.laparoscope
    tips, bases, radii, heights = lap.initial_points()
    oi = orc.OracleInstrument(np.concatenate([lap.stored_transforms()] * n), np.tile(heights, (n, 1)),
                              np.tile(tips, (n, 1)), np.tile(bases, (n, 1)), np.float32)
    o = orc.OracleSim([mesh], [0] * n, np.float32, tips=oi.tips.copy())
    o64 = orc.OracleSim([mesh], [0] * n, np.float64, tips=oi.tips.copy())
    cfg = orc.StepConfig(**ORACLE)
    integ = pkg.SimIntegratorDO(gpu)
    diag = mesh.bbox_diagonal()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(gpu)

    def check(what, positions=False):
        torch.cuda.synchronize()
        t = model.laparoscopeTip.cpu().numpy().reshape(-1, 3)
        e_tip = float(np.abs(t - oi.tips).max())
        assert e_tip <= 5e-5, (what, e_tip)
        assert model.laparoscopeDragCutHeat.cpu().numpy().tolist() == oi.dch.tolist(), what
        x = model.vertex().cpu().numpy().reshape(n, -1, 3)
        assert all(np.array_equal(x[0], x[e]) for e in range(1, n)), what
        e_x = float(np.abs(x.reshape(-1, 3) - o.x).max())
        e_64 = float(np.abs(o.x - o64.x).max())
        print(f"\n[liver {what}] |gpu - o32| {e_x / diag:.2e} diag, |o32 - o64| {e_64 / diag:.2e} diag", end="")
        if positions:
            assert e_64 <= 5e-6 * diag, (what, e_64 / diag)        # inside the reproducible horizon
            assert e_x <= 1e-5 * diag, (what, e_x / diag)
        return e_tip, e_x, e_64

    # testLiverRetraction.py:51-58: rotation centre, then the effector to (0, 10, 4)
    model.setRotationCenter(dev(np.tile(RCM, (n, 1))), dev(np.tile(RCM_ROT, (n, 1))))
    oi.set_rcm(np.tile(RCM, (n, 1)), np.tile(RCM_ROT, (n, 1)))
    shift = np.asarray(EFFECTOR, np.float32) - model.getLaparoscopePositionsTensor().cpu().numpy()
    model.applyCartesianActions(dev(shift))
    oi.cartesian(shift)
    check("set-up", positions=True)
    assert np.allclose(oi.tips.reshape(n, 4, 3)[:, 3], EFFECTOR, atol=1e-4)

    def frame():
        integ.stepModel(model)
        for oo in (o, o64):
            oo.tips = oi.tips.astype(oo.tips.dtype)
            oo.step(cfg, 1)

    # approach the nearest free vertex, 0.1 per frame (actionStrength of the test)
    free = mesh.inv_mass > 0
    d = np.linalg.norm(mesh.verts - np.asarray(EFFECTOR), axis=1)
    d[~free] = np.inf
    step = mesh.verts[np.argmin(d)] - np.asarray(EFFECTOR)
    step = (step / np.linalg.norm(step) * 0.1).astype(np.float32)
    for f in range(13):
        a = np.tile(step, (n, 1))
        model.applyCartesianActions(dev(a))
        oi.cartesian(a)
        frame()
        check(f"approach {f}", positions=f < HORIZON)
    # close the jaws (-0.5 rad per call from the scene's 80.5 degrees): the
    # grasp switches on at the first call that reads an angle below 43 degrees
    close = np.zeros((n, 5), np.float32)
    close[:, 4] = -0.5
    for _ in range(3):
        model.applyActions(dev(close))
        oi.actions(close)
        orc.call("oracle_grasp_select", np.float32, n * mesh.num_vertices, o.x, o.v2e, oi.tips, o.drag, 1.0,
                 oi.dch, oi.dchu)
        torch.cuda.synchronize()
        assert np.array_equal(model.drag.cpu().numpy(), o.drag)
    o64.drag[:] = o.drag
    grasped = int(o.drag.sum()) // n
    assert grasped > 0 and oi.dch[0] == 1
    check("grasp")
    # lift (retraction), 0.1 per frame
    y0 = o.x.reshape(n, -1, 3)[0][o.drag[:mesh.num_vertices] > 0][:, 1].mean()
    for f in range(10):
        a = np.tile([0.0, 0.1, 0.0], (n, 1)).astype(np.float32)
        model.applyCartesianActions(dev(a))
        oi.cartesian(a)
        frame()
        e_tip, e_x, e_64 = check(f"lift {f}")
    y1 = o.x.reshape(n, -1, 3)[0][o.drag[:mesh.num_vertices] > 0][:, 1].mean()
    print(f"\n[liver] grasped {grasped} vertices per env, lifted by {y1 - y0:.3f}; last |tip err| {e_tip:.2e}, "
          f"|gpu - o32| {e_x / diag:.2e} diag, |o32 - o64| {e_64 / diag:.2e} diag")
    assert y1 > y0 + 0.2
    # workspace-limited move: the drag point is above workspaceHigh.y, so the y
    # component is dropped and the env flagged (applyCartesianActionsInWorkspace)
    a = np.tile([0.1, -0.1, 0.0], (n, 1)).astype(np.float32)
    model.applyCartesianActionsInWorkspace(dev(a))
    oi.cartesian(a, np.array(WS_LO, np.float32), np.array(WS_HI, np.float32))
    torch.cuda.synchronize()
    assert model.getLaparoscopeWorkspaceViolation().cpu().numpy().tolist() == oi.violation.tolist() == [1] * n
    check("workspace move")


def test_stochastic_reset_and_forced_clamp_vertex(pkg, gpu):
    """resetStochasticModel (modelDO.py:1624-1645): positions and instrument
    back to their initial state, clamp state cleared, then re-aimed at the
    given rotation centre; velocities are left as they are (as in the
    reference).  forceLaparoscopeClampVertex (modelDO.py:1533-1544): grasp
    weight of one local vertex in the selected envs only."""
    n = 3
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(gpu)
    model = pkg.SimModelDO.from_usd(FIXTURE, n, device=gpu, **REF)
    integ = pkg.SimIntegratorDO(gpu)
    model.applyCartesianActions(dev(np.tile([0.3, -0.4, 0.2], (n, 1))))
    close = np.zeros((n, 5), np.float32)
    close[:, 4] = -0.5
    model.applyActions(dev(close))
    for _ in range(3):
        integ.stepModel(model)
    free = int(np.nonzero(model.scene.meshes[0].inv_mass > 0)[0][0])
    model.forceLaparoscopeClampVertex(free, torch.tensor([1, 0, 1]), 1.0)
    torch.cuda.synchronize()
    d = model.drag.cpu().numpy()
    vb = model.env_vbase
    assert d[vb[0] + free] == 1.0 and d[vb[2] + free] == 1.0
    with pytest.raises(IndexError):
        model.forceLaparoscopeClampVertex(10 ** 6, torch.tensor([1, 0, 0]))
    v_before = model.v.clone()
    assert not torch.equal(model.x, model.initialVertex)
    model.resetStochasticModel(dev(np.tile(RCM, (n, 1))), dev(np.tile(RCM_ROT, (n, 1))))
    fresh = pkg.SimModelDO.from_usd(FIXTURE, n, device=gpu, **REF)
    fresh.setRotationCenter(dev(np.tile(RCM, (n, 1))), dev(np.tile(RCM_ROT, (n, 1))))
    torch.cuda.synchronize()
    assert torch.equal(model.x, model.initialVertex) and torch.equal(model.v, v_before)
    assert torch.equal(model.laparoscopeXForm, fresh.laparoscopeXForm)
    assert torch.equal(model.laparoscopeTip, fresh.laparoscopeTip)
    assert torch.equal(model.laparoscopeBase, fresh.laparoscopeBase)
    assert int(model.laparoscopeDragCutHeat.abs().sum()) == 0


def test_remap_render_points(pkg, orc, gpu):
    """ffsrl_remap (remapAToBKernel, utils.py:32-42) on the scene's own
    render map (tests/golden/liver_render_map.npz, mapping:pointToVertex of
    liverRetraction.usd): at rest the render points come back within 2e-5
    (float32 world transform, coordinates up to 12); after two frames every
    env's remapped points equal the oracle's remap of the GPU positions bit
    for bit (a float4 copy, w included)."""
    r = np.load(os.path.join(os.path.dirname(FIXTURE), "liver_render_map.npz"))
    points, pairs = r["render_points"], r["point_to_vertex"].astype(np.int32)
    n = 2
    model = pkg.SimModelDO.from_usd(FIXTURE, n, device=gpu, **REF)
    integ = pkg.SimIntegratorDO(gpu)
    P = points.shape[0]
    assert pairs[:, 0].max() < P and pairs[:, 1].max() < model.meshes[0].num_vertices

    def remap_all():
        out = torch.zeros((n, P, 4), dtype=torch.float32, device=gpu)
        for e in range(n):
            m = pairs.copy()
            m[:, 1] += int(model.env_vbase[e])
            dm = torch.from_numpy(m.reshape(-1)).to(gpu)
            pkg.lib.ffsrl_remap(out[e].data_ptr(), model.x.data_ptr(), dm.data_ptr(), m.shape[0], model.stream())
        torch.cuda.synchronize()
        return out.cpu().numpy()

    rest = remap_all()
    assert np.abs(rest[..., :3] - points[None]).max() <= 2e-5
    for _ in range(2):
        integ.stepModel(model)
    got = remap_all()
    x = model.x.cpu().numpy()
    for e in range(n):
        m = pairs.copy()
        m[:, 1] += int(model.env_vbase[e])
        want = orc.remap(np.zeros((P, 4), np.float32), x, m)
        assert np.array_equal(got[e], want), e
    assert not np.array_equal(got[0], rest[0])                 # the frames moved the scene
