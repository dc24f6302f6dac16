"""The reference's launch model (ffsrl_frame_do, SimIntegratorDO.stepModelDO):
every plain kernel launched once per substep and iteration over all
environments (an env is found from the prefix sums of the envs' item
counts), and the refit as one launch per refit group over all
environments.  Against the per-env plain form, which runs the same kernels
one env at a time (their float atomics add in another order): positions
within 1e-6 x the mesh diagonal; against the float32 oracle within 1e-5 x
diag over the 3 frames where the scheme is reproducible; BVH bounds equal
the oracle's refit of the same positions bit for bit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MIXED = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
             environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
             volumeComplianceXPBD=0.0, neoHookean=True, globalVolumeCompliance=1.0 / 5e4,
             globalDeviatoricCompliance=1.0 / 1e4, globalKsDrag=1.0)
MIXED_ORACLE = dict(substeps=10, iterations=1, fps=60, gravity=(0.0, -9.81, 0.0), damping=0.1, dist_mode=2,
                    vol_mode=2, nh_on=1, nh_vol_c=1.0 / 5e4, nh_dev_c=1.0 / 1e4, dist_compliance=1e-6,
                    vol_compliance=0.0, drag_ks=1.0)


def _meshes(pkg, n, lo, hi, seed):
    sc = pkg.scenes
    specs = sc.mixed_sizes(n, lo, hi, seed=seed)[0]
    return [sc.solid(shape, int(t), seed=seed + i) for i, (shape, t) in enumerate(specs)]


def _models(pkg, gpu, meshes, kw, k=2, grasp=True):
    sc = pkg.scenes
    tips = np.stack([np.repeat((sc.top_point(m) + np.array([0.0, 0.02, 0.0], np.float32))[None], 4, axis=0)
                     for m in meshes]).astype(np.float32)
    out = []
    for _ in range(k):
        m = pkg.SimModelDO(meshes, list(range(len(meshes))), device=gpu, tips=tips, **kw)
        if grasp:
            for e in range(1, len(meshes), 2):
                m.set_grasp(e, sc.nearest_surface_vertices(meshes[e], tips[e, 2], 1))
        out.append(m)
    return out, tips


def test_one_launch_equals_plain_and_oracle(pkg, orc, gpu):
    meshes = _meshes(pkg, 6, 1000, 8000, seed=61)
    (a, b), tips = _models(pkg, gpu, meshes, MIXED)
    integ = pkg.SimIntegratorDO(gpu)
    o32 = orc.OracleSim(meshes, list(range(len(meshes))), np.float32, tips=tips)
    for e in range(1, len(meshes), 2):
        g = pkg.scenes.nearest_surface_vertices(meshes[e], tips[e, 2], 1)
        o32.drag[np.asarray(g) + o32.vbase[e]] = 1.0
    cfg = orc.StepConfig(**MIXED_ORACLE)
    for f in range(3):
        integ.stepModelDO(a, refit=True)
        integ.stepModelPlain(b)
        pkg.lib.ffsrl_bvh_refit_plain(b.batch, 0, b.numEnvs, b.stream())
        o32.step(cfg, 1)
    torch.cuda.synchronize()
    xa, xb = a.vertex().cpu().numpy(), b.vertex().cpu().numpy()
    lo, hi = a.bvh_lo.cpu().numpy(), a.bvh_hi.cpu().numpy()
    for e, m in enumerate(meshes):
        s = a.env_slice(e)
        diag = m.bbox_diagonal()
        e_plain = float(np.abs(xa[s] - xb[s]).max()) / diag
        e_o32 = float(np.abs(xa[s] - o32.x[s]).max()) / diag
        print(f"\n[one launch] env {e} ({m.num_tets} tets): |do-plain| {e_plain:.2e} diag, |do-o32| {e_o32:.2e} diag")
        assert e_plain <= 1e-6 and e_o32 <= 1e-5
        t = a.topo[e]
        bvh = {"left": t.left, "first": t.first, "count": t.count, "level": t.level, "tri_map": t.tri_map}
        mn, mx = orc.refit(bvh, m.tris, xa[s])
        nb = int(a.env_nbase[e])
        assert np.array_equal(lo[nb:nb + t.num_nodes, :3], mn) and np.array_equal(hi[nb:nb + t.num_nodes, :3], mx)


def test_one_launch_carried_lambdas(pkg, gpu):
    kw = dict(MIXED, neoHookean=False, simConstraintsSteps=2, simSubsteps=5)
    meshes = _meshes(pkg, 4, 1000, 6000, seed=67)
    (a, b), _ = _models(pkg, gpu, meshes, kw)
    assert a.lam_t is not None
    integ = pkg.SimIntegratorDO(gpu)
    for f in range(2):
        integ.stepModelDO(a)
        integ.stepModelPlain(b)
    torch.cuda.synchronize()
    for e, m in enumerate(meshes):
        s = a.env_slice(e)
        err = float((a.vertex()[s] - b.vertex()[s]).abs().max()) / m.bbox_diagonal()
        assert err <= 1e-6, (e, err)
    assert bool(torch.isfinite(a.lam_t).all()) and float(a.lam_t.abs().max()) > 0.0


def test_one_launch_collision(pkg, gpu):
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 3000, seed=71 + i) for i in range(3)]
    R = 0.12
    bases, tips = [], []
    for m in meshes:
        q = sc.top_point(m).astype(np.float64)
        c0 = q + np.array([0.0, 0.7 * R, 0.0])
        bases.append(np.stack([c0 - [0, R, 0], q + [-0.15, 0.0, 0.0], q + [0, 3.0, 0], q + [0, 0.5, 0]]))
        tips.append(np.stack([c0 + [0, 0.5 + R, 0], q + [-0.15, 0.6, 0.0], q + [0, 3.5, 0], q + [0, 0.5, 0]]))
    f32 = lambda v: np.asarray(v, np.float32)
    kw = dict(MIXED, neoHookean=False)
    a, b = [pkg.SimModelDO(meshes, [0, 1, 2], device=gpu, tips=f32(tips), capsuleBases=f32(bases),
                           capsuleRadii=f32([[R, R, 0.05, 0.0]] * 3), collision=True, globalKsContact=1.0, **kw)
            for _ in range(2)]
    integ = pkg.SimIntegratorDO(gpu)
    d = torch.tensor([[0.0, -0.004, 0.0]] * 3, dtype=torch.float32, device=gpu)
    for f in range(2):
        for m in (a, b):
            pkg.lib.ffsrl_move_tools(m.batch, d.data_ptr(), None, None, m.stream())
        integ.stepModelDO(a)
        integ.stepModelPlain(b)
    torch.cuda.synchronize()
    assert int(a.laparoscopeInCollision.sum()) > 0
    assert torch.equal(a.laparoscopeInCollision, b.laparoscopeInCollision)
    for e, m in enumerate(meshes):
        s = a.env_slice(e)
        err = float((a.vertex()[s] - b.vertex()[s]).abs().max()) / m.bbox_diagonal()
        assert err <= 1e-6, (e, err)
