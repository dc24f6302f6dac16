"""GPU parity of the instrument collision (collideLaparoscopeDO,
integrator.py:981-1050) against the oracle: capsules pressing down into
seeded ellipsoids, batched and plain forms.

Tolerances: vertex positions within 1e-5 x the mesh bounding-box diagonal of
the fp32 oracle (as for the other physics parity tests), at horizons where
the fp32 and fp64 oracles themselves agree to that bound; the per-env contact
counts may differ where a triangle sits within float rounding of the capsule
surface (distance == radius), so they match to 1% (and at least 2 counts).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DV = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
          environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
          volumeComplianceXPBD=0.0)
DV_ORACLE = dict(substeps=10, iterations=1, fps=60, gravity=(0.0, -9.81, 0.0), damping=0.1, dist_mode=2, vol_mode=2,
                 dist_compliance=1e-6, vol_compliance=0.0)
RADIUS = 0.12


def _scene(pkg, n_env, tets, seed, depth=0.3, side=0.15, RADIUS=RADIUS):
    """Per env: capsule 0 vertical over the highest surface vertex, its lower
    sphere `depth` radii into the body; capsule 1 tilted 45 degrees, placed the
    same way `side` to the -x side; capsule 2 far above; point 3 the drag
    point (radius 0)."""
    sc = pkg.scenes
    meshes, bases, tips, radii = [], [], [], []
    for i in range(n_env):
        m = sc.solid("ellipsoid", tets, seed=seed + i)
        q = sc.top_point(m).astype(np.float64)
        up = np.array([0.0, 1.0, 0.0])
        tilt = np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
        c0 = q + up * ((1.0 - depth) * RADIUS)             # lower sphere centre of capsule 0
        c1 = q + np.array([-side, 0.0, 0.0]) + up * ((1.0 - depth) * RADIUS)
        b = np.stack([c0 - up * RADIUS, c1 - tilt * RADIUS, q + [0.0, 3.0, 0.0], q + [0.0, 0.5, 0.0]])
        t = np.stack([c0 + up * (0.5 + RADIUS), c1 + tilt * (0.5 + RADIUS), q + [0.0, 3.5, 0.0],
                      q + [0.0, 0.5, 0.0]])
        meshes.append(m)
        bases.append(b)
        tips.append(t)
        radii.append([RADIUS, RADIUS, 0.05, 0.0])
    f32 = lambda a: np.asarray(a, np.float32)
    return meshes, f32(bases), f32(tips), f32(radii)


C0 = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, 0.0, 0.0),
          environmentVelocityDampening=0.1, distance=None, volume=None)
C0_ORACLE = dict(substeps=10, iterations=1, fps=60, gravity=(0.0, 0.0, 0.0), damping=0.1, dist_mode=0, vol_mode=0)


def _press(pkg, orc, gpu, form, kw, okw, k, frames, depth, n_env=3, tets=5000, seed=21, o64=True):
    """Capsules pressing 4 mm per frame: the model after `frames` frames and
    the fp32 (and fp64) oracles run on the same inputs."""
    meshes, bases, tips, radii = _scene(pkg, n_env, tets, seed, depth)
    envs = list(range(n_env))
    model = pkg.SimModelDO(meshes, envs, device=gpu, tips=tips, capsuleBases=bases, capsuleRadii=radii,
                           collision=True, globalKsContact=k, **kw)
    oracles = [orc.OracleSim(meshes, envs, dt, tips=tips, caps_base=bases, caps_radius=radii)
               for dt in ((np.float32, np.float64) if o64 else (np.float32,))]
    cfg = orc.StepConfig(collide=1, contact_k=k, **okw)
    integ = pkg.SimIntegratorDO(gpu, block=512 if form == "batched512" else 1024)
    delta = np.tile(np.array([[0.0, -0.004, 0.0]], np.float32), (n_env, 1))
    dt = torch.from_numpy(delta).to(gpu)
    step = np.repeat(delta, 4, axis=0)
    for _ in range(frames):
        pkg.lib.ffsrl_move_tools(model.batch, dt.data_ptr(), None, None, model.stream())
        if form == "plain":
            integ.stepModelPlain(model)
        else:
            integ.stepModel(model)
        for o in oracles:
            o.tips = (o.tips + step).astype(o.tips.dtype)
            o.caps_base = (o.caps_base + step).astype(o.caps_base.dtype)
            o.step(cfg, 1)
    torch.cuda.synchronize()
    # the capsule bases moved with their tips
    assert np.array_equal(model.laparoscopeTip.cpu().numpy().reshape(-1, 3), oracles[0].tips)
    assert np.array_equal(model.laparoscopeBase.cpu().numpy().reshape(-1, 3), oracles[0].caps_base)
    return meshes, model, oracles


def _check(name, meshes, model, o32, o64):
    x = model.vertex().cpu().numpy()
    for i in range(len(meshes)):
        s = slice(o32.vbase[i], o32.vbase[i + 1])
        diag = meshes[i].bbox_diagonal()
        e32 = float(np.abs(x[s] - o32.x[s]).max())
        e64 = float(np.abs(o32.x[s] - o64.x[s]).max())
        print(f"\n[{name} env{i}] diag={diag:.3f} |gpu-o32|={e32:.3e} ({e32 / diag:.2e} diag) |o32-o64|={e64:.3e}")
        assert e32 <= 1e-5 * diag
    got = model.getLaparoscopeInCollision().cpu().numpy()
    want = o32.in_collision
    print(f"[{name}] contacts gpu {got.tolist()} oracle {want.tolist()} fp64 oracle {o64.in_collision.tolist()}")
    assert (want > 5).all()
    assert (np.abs(got - want) <= np.maximum(2, want // 100)).all()


@pytest.mark.parametrize("form", ["batched", "batched512", "plain"])
def test_collision_pass(pkg, orc, gpu, form):
    """The collision alone (no constraints, no gravity): capsules half a
    radius into the bodies, pressing for 4 frames (40 passes)."""
    meshes, model, (o32, o64) = _press(pkg, orc, gpu, form, C0, C0_ORACLE, 1.0, 4, 0.5)
    _check(f"collision pass {form}", meshes, model, o32, o64)


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_collision_xpbd(pkg, orc, gpu, form):
    """Collision inside the XPBD distance + volume scheme with gravity.  The
    press is ill-conditioned in the reference's formulation (the fp32 and
    fp64 oracles part by > 1e-5 x diag after 3 frames with contact stiffness
    0.25, after 2 with 1.0), so parity is checked at 2 frames."""
    meshes, model, (o32, o64) = _press(pkg, orc, gpu, form, DV, DV_ORACLE, 0.25, 2, 0.3)
    _check(f"collision xpbd {form}", meshes, model, o32, o64)


def test_collision_press_30_frames(pkg, orc, gpu):
    """30 frames of the press (12 cm) in the XPBD scheme, past the parity
    horizon: the positions stay finite, the bodies are pressed down against
    a run without the instrument, contacts keep counting and the counters
    clear."""
    meshes, model, (o32,) = _press(pkg, orc, gpu, "batched", DV, DV_ORACLE, 1.0, 30, 0.3, o64=False)
    x = model.vertex().cpu().numpy()
    assert np.isfinite(x).all()
    free = orc.OracleSim(meshes, list(range(len(meshes))), np.float32)
    free.step(orc.StepConfig(**DV_ORACLE), 30)
    got = model.getLaparoscopeInCollision().cpu().numpy()
    print(f"\n[collision 30 frames] contacts gpu {got.tolist()} oracle {o32.in_collision.tolist()}")
    assert (got > 100).all()
    for i in range(len(meshes)):
        s = slice(o32.vbase[i], o32.vbase[i + 1])
        top = int(np.argmax(meshes[i].verts[:, 1]))
        print(f"[collision 30 frames env{i}] top y gpu {x[s][top, 1]:.4f} oracle {o32.x[s][top, 1]:.4f} "
              f"free {free.x[s][top, 1]:.4f}")
        assert x[s][top, 1] < free.x[s][top, 1] - 0.05
    model.resetCollisionInfo()
    assert int(model.getLaparoscopeInCollision().sum()) == 0


def test_collision_off_changes_nothing(pkg, gpu):
    """collision=False with capsules present runs the frame without the pass:
    positions equal a model built without capsules, and no contacts count."""
    meshes, bases, tips, radii = _scene(pkg, 2, 2500, 5)
    a = pkg.SimModelDO(meshes, [0, 1], device=gpu, tips=tips, capsuleBases=bases, capsuleRadii=radii, **DV)
    b = pkg.SimModelDO(meshes, [0, 1], device=gpu, tips=tips, **DV)
    integ = pkg.SimIntegratorDO(gpu)
    for _ in range(5):
        integ.stepModel(a)
        integ.stepModel(b)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x)
    assert int(a.getLaparoscopeInCollision().sum()) == 0


def test_reset_restores_capsule_bases(pkg, gpu):
    meshes, bases, tips, radii = _scene(pkg, 2, 2500, 5)
    model = pkg.SimModelDO(meshes, [0, 1], device=gpu, tips=tips, capsuleBases=bases, capsuleRadii=radii,
                           collision=True, **DV)
    dt = torch.tensor([[0.01, -0.02, 0.0], [0.0, 0.0, 0.03]], dtype=torch.float32, device=gpu)
    pkg.lib.ffsrl_move_tools(model.batch, dt.data_ptr(), None, None, model.stream())
    torch.cuda.synchronize()
    moved = model.laparoscopeBase.cpu().numpy()
    assert np.allclose(moved - bases, np.repeat(dt.cpu().numpy()[:, None, :], 4, axis=1), atol=1e-6)
    model.resetFixedModel([0])
    assert np.array_equal(model.laparoscopeBase[0].cpu().numpy(), bases[0])
    assert np.array_equal(model.laparoscopeBase[1].cpu().numpy(), moved[1])


@pytest.mark.parametrize("radius,ks,saturates", [(0.12, 1.0, False), (0.6, 4.0, None), (0.6, 16.0, True)])
def test_deep_contact_is_exact_or_flagged(pkg, gpu, radius, ks, saturates):
    """The batched frame's fixed-point terms hold +-(2^24 - 1) quanta, D/8 ..
    D/4 of the mesh diagonal D (0.5 on these 2.5-diagonal meshes), while a
    collision response reaches up to 3 r ks (r the capsule radius): plen <= r
    and the barycentric weights' squared norm is >= 1/3.  A moderate press (r 0.12, ks 1) stays in range: the batched
    frame equals the plain form within 1e-5 x diag and reports nothing.  A
    deep press with a large stiff capsule (r 0.6, ks 16) produces terms far
    beyond the range: the batched frame clamps them and reports
    STATUS_SATURATED for the env, where it no longer equals the plain form
    (which sums floats).  In between (r 0.6, ks 4) either holds."""
    meshes, bases, tips, radii = _scene(pkg, 2, 5000, 31, depth=0.5, RADIUS=radius)
    kw = dict(C0, collision=True, globalKsContact=ks, tips=tips, capsuleBases=bases, capsuleRadii=radii)
    models = {f: pkg.SimModelDO(meshes, [0, 1], device=gpu, **kw) for f in ("batched", "plain")}
    integ = pkg.SimIntegratorDO(gpu)
    integ.stepModel(models["batched"])
    integ.stepModelPlain(models["plain"])
    torch.cuda.synchronize()
    st = models["batched"].frame_status().cpu().numpy()
    xb, xp = models["batched"].vertex().cpu().numpy(), models["plain"].vertex().cpu().numpy()
    for e, m in enumerate(meshes):
        s = models["batched"].env_slice(e)
        err = float(np.abs(xb[s] - xp[s]).max()) / m.bbox_diagonal()
        moved = float(np.nanmax(np.abs(xb[s] - m.verts)))                # the plain form may not stay finite
        print(f"\n[saturation r={radius} ks={ks}] env {e}: status {int(st[e])}, |batched-plain| {err:.2e} diag, "
              f"largest move {moved:.3f}")
        assert moved > 1e-3                                             # the capsule pressed the body
        if saturates:
            assert st[e] & pkg._lib.STATUS_SATURATED
        elif saturates is not None:
            assert st[e] == 0 and err <= 1e-5
        assert st[e] != 0 or err <= 1e-5                               # exact, or reported
