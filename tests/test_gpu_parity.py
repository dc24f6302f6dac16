"""GPU parity: the HIP path through the C ABI against the oracle.

Tolerances (from the path's definition of done):
  * vertex positions after N frames within 1e-5 x the mesh bounding-box
    diagonal of the fp32 oracle; the errors of both against the fp64 oracle
    are printed;
  * BVH node bounds and ray-hit triangle ids bit-exact on the same positions;
  * depth bit-exact on the same positions (it is a function of the hit t).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

C1 = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
          environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
          volumeComplianceXPBD=0.0)
C1_ORACLE = dict(substeps=10, iterations=1, fps=60, gravity=(0.0, -9.81, 0.0), damping=0.1, dist_mode=2, vol_mode=2,
                 dist_compliance=1e-6, vol_compliance=0.0)
NH = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
          environmentVelocityDampening=0.1, distance=None, volume=None, neoHookean=True,
          globalVolumeCompliance=1.0 / 5e4, globalDeviatoricCompliance=1.0 / 1e4, globalKsDrag=1.0)
NH_ORACLE = dict(substeps=10, iterations=1, fps=60, gravity=(0.0, -9.81, 0.0), damping=0.1, nh_on=1,
                 nh_vol_c=1.0 / 5e4, nh_dev_c=1.0 / 1e4, drag_ks=1.0)


def _report(name, gpu_x, o32, o64, diag):
    e32 = float(np.abs(gpu_x - o32).max())
    eg64 = float(np.abs(gpu_x - o64).max())
    e3264 = float(np.abs(o32 - o64).max())
    print(f"\n[{name}] diag={diag:.4f} |gpu-o32|={e32:.3e} ({e32 / diag:.2e} diag) "
          f"|gpu-o64|={eg64:.3e} |o32-o64|={e3264:.3e}")
    return e32


def _c1_model(pkg, n=10, frames_velocity_seed=7):
    mesh = pkg.scenes.cube_grid(n)
    rng = np.random.default_rng(frames_velocity_seed)
    v0 = rng.uniform(-0.01, 0.01, size=(mesh.num_vertices, 3)).astype(np.float32)
    return mesh, v0


@pytest.mark.parametrize("form", ["batched", "batched512", "batched256", "plain"])
def test_c1_cube_100_frames(pkg, orc, gpu, form):
    mesh, v0 = _c1_model(pkg)
    model = pkg.SimModelDO([mesh], [0], device=gpu, initialVelocity=v0, **C1)
    integ = pkg.SimIntegratorDO(gpu, block={"batched512": 512, "batched256": 256}.get(form, 1024))
    frames = 100
    for _ in range(frames):
        if form.startswith("batched"):
            integ.stepModel(model)
        else:
            integ.stepModelPlain(model)
    torch.cuda.synchronize()
    x = model.vertex().cpu().numpy()
    cfg = orc.StepConfig(**C1_ORACLE)
    o32 = orc.OracleSim([mesh], [0], np.float32, initial_velocity=v0)
    o64 = orc.OracleSim([mesh], [0], np.float64, initial_velocity=v0)
    o32.step(cfg, frames)
    o64.step(cfg, frames)
    diag = mesh.bbox_diagonal()
    e = _report(f"C1 {form}", x, o32.x, o64.x, diag)
    assert np.isfinite(x).all()
    assert e <= 1e-5 * diag


def _bvh_dict(topo):
    return {"left": topo.left, "first": topo.first, "count": topo.count, "level": topo.level,
            "tri_map": topo.tri_map}


@pytest.mark.parametrize("traversal", [1, 2])
def test_refit_and_depth_bit_exact(pkg, orc, gpu, traversal):
    """C1 camera: 128x128 depth (256x256 rays) after 20 frames with the refit
    fused into the frame; bounds, per-ray hits and depth equal the oracle's,
    with one ray per lane (1) and with 2x2-ray quads per lane (2)."""
    mesh, v0 = _c1_model(pkg)
    model = pkg.SimModelDO([mesh], [0], device=gpu, initialVelocity=v0, **C1)
    integ = pkg.SimIntegratorDO(gpu)
    for _ in range(20):
        integ.stepModel(model, refit=True)
    ren = pkg.RaycastRendererDO(gpu, model, cameraPos=(2.2, 1.8, 2.6), resolution=128, traversal=traversal)
    ren.lookAt((0.5, 0.45, 0.5))
    hit = torch.full((4 * 128 * 128,), -7, dtype=torch.int32, device=gpu)
    thit = torch.zeros(4 * 128 * 128, dtype=torch.float32, device=gpu)
    depth = ren.render_depth(hit=hit, thit=thit)
    torch.cuda.synchronize()
    x = model.vertex().cpu().numpy()
    bvh = _bvh_dict(model.topo[0])
    mn, mx = orc.refit(bvh, mesh.tris, x)
    assert np.array_equal(model.bvh_lo[:, :3].cpu().numpy(), mn)
    assert np.array_equal(model.bvh_hi[:, :3].cpu().numpy(), mx)
    d_ref, h_ref, t_ref, ms = orc.render_depth(128, 128, np.asarray(ren.camera.origin[:], np.float32),
                                               ren.dirs.cpu().numpy(), bvh, mn, mx, mesh.tris, x)
    h = hit.cpu().numpy()
    print(f"\n[render C1] hits {(h_ref >= 0).sum()} / {h_ref.size}, max stack {ms}")
    assert (h_ref >= 0).sum() > 0.2 * h_ref.size
    assert np.array_equal(h, h_ref)
    assert np.array_equal(thit.cpu().numpy(), t_ref)
    assert np.array_equal(depth[0].cpu().numpy(), d_ref)


def _edge_rays(n, rng):
    """Ray directions that exercise every branch of the slab test: exact axis
    directions, directions with one zero component (infinite reciprocals,
    and NaN products where a box face lies in the origin's plane),
    components below and at the 1e-30 limit of the fast path, and random
    directions towards the unit cube."""
    d = []
    for ax in np.eye(3):
        d += [-ax, ax]
    for a in np.linspace(-0.9, 0.9, 40):
        d += [(0.0, a, -1.0), (a, 0.0, -1.0), (a, -1.0, 0.0)]
    for tiny in (1e-38, 1e-35, -1e-33, 5e-31, 1e-30, -1e-30, 3e-30, 1e-25):
        for a in (-0.4, 0.1, 0.3):
            d += [(tiny, a, -1.0), (a, tiny, -1.0), (a, -1.0, tiny)]
    d = np.array(d, np.float64)
    rest = rng.normal(size=(n - len(d), 3))
    rest[:, 2] = -np.abs(rest[:, 2]) - 1.0
    d = np.concatenate([d, rest])
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    out = np.zeros((n, 4), np.float32)
    out[:, :3] = d.astype(np.float32)
    return out


@pytest.mark.parametrize("traversal", [1, 2])
def test_depth_edge_rays_bit_exact(pkg, orc, gpu, traversal):
    """Depth, hits and t for rays with zero, tiny and axis-aligned direction
    components, from origins that lie in planes of the mesh's grid, equal the
    oracle's (the division form of the reference) bit for bit.  In the quad
    traversal (2) the four rays of a pixel point in unrelated directions
    here, so their reach masks part at the root."""
    mesh, _ = _c1_model(pkg)
    model = pkg.SimModelDO([mesh], [0], device=gpu, **C1)
    pkg.SimBVH(gpu, model)
    x = model.vertex().cpu().numpy()
    xs = np.unique(x[:, 0])
    ys = np.unique(x[:, 1])
    bvh = _bvh_dict(model.topo[0])
    mn, mx = orc.refit(bvh, mesh.tris, x)
    rng = np.random.default_rng(5)
    ren = pkg.RaycastRendererDO(gpu, model, cameraPos=(0.0, 0.5, 2.5), resolution=16, traversal=traversal)
    for origin in ((xs[0], ys[4], 2.5), (xs[3], ys[0], 2.5), (xs[5], ys[5], 2.5)):
        ren.camera.origin[:] = tuple(float(c) for c in np.asarray(origin, np.float32))
        dirs = _edge_rays(32 * 32, rng)
        ren.dirs.copy_(torch.from_numpy(dirs))
        hit = torch.full((4 * 16 * 16,), -7, dtype=torch.int32, device=gpu)
        thit = torch.zeros(4 * 16 * 16, dtype=torch.float32, device=gpu)
        depth = ren.render_depth(hit=hit, thit=thit)
        torch.cuda.synchronize()
        d_ref, h_ref, t_ref, _ = orc.render_depth(16, 16, np.asarray(ren.camera.origin[:], np.float32), dirs,
                                                  bvh, mn, mx, mesh.tris, x)
        assert (h_ref >= 0).sum() > 100
        assert np.array_equal(hit.cpu().numpy(), h_ref), origin
        assert np.array_equal(thit.cpu().numpy(), t_ref), origin
        assert np.array_equal(depth[0].cpu().numpy(), d_ref), origin


def test_refit_fused_equals_standalone(pkg, gpu):
    mesh, v0 = _c1_model(pkg, 8)
    model = pkg.SimModelDO([mesh], [0], device=gpu, initialVelocity=v0, **C1)
    integ = pkg.SimIntegratorDO(gpu)
    integ.stepModel(model, refit=True)
    lo, hi = model.bvh_lo.clone(), model.bvh_hi.clone()
    model.bvh_lo.zero_()
    model.bvh_hi.zero_()
    pkg.SimBVH(gpu, model)                     # refit on construction
    assert torch.equal(lo, model.bvh_lo) and torch.equal(hi, model.bvh_hi)
    model.bvh_lo.zero_()
    model.bvh_hi.zero_()
    pkg.SimBVH(gpu, model).refitBVH(plain=True)
    assert torch.equal(lo, model.bvh_lo) and torch.equal(hi, model.bvh_hi)


def _grasp_scene(pkg, n_env, tets, seed):
    sc = pkg.scenes
    meshes, tips, grasp = [], [], []
    for i in range(n_env):
        m = sc.solid("ellipsoid", tets, seed=seed + i)
        meshes.append(m)
        tip = sc.top_point(m) + np.array([0.0, 0.02, 0.0], np.float32)
        tips.append(np.repeat(tip[None], 4, axis=0))
        grasp.append(sc.nearest_surface_vertices(m, tip, 8))
    return meshes, np.stack(tips), grasp


NH_DV = dict(NH, distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0)
NH_DV_ORACLE = dict(NH_ORACLE, dist_mode=2, vol_mode=2, dist_compliance=1e-6, vol_compliance=0.0)


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("case", ["nh_grasp8_2frames", "nh_dv_grasp1_30frames"])
def test_c2_neohookean_grasp_random_walk(pkg, orc, gpu, form, case):
    """Config-2 scheme at reduced size: Neo-Hookean (mu 1e4, lambda 5e4) with
    grasped surface vertices pulled toward a tool tip on a seeded random walk
    (sigma 2 mm).  Horizons follow the scheme's conditioning: with 8 grasped
    vertices pulled to one point (dragConstraintsDO) the NH-only body
    collapses under gravity and the fp32 and fp64 oracles themselves drift
    apart by > 1e-5 x diag after ~3 frames, so that case runs 2 frames; with
    distance + volume constraints and one grasped vertex the scheme is well
    conditioned and runs 30 frames."""
    if case == "nh_grasp8_2frames":
        n_env, frames, k, kw, okw = 3, 2, 8, NH, NH_ORACLE
    else:
        n_env, frames, k, kw, okw = 3, 30, 1, NH_DV, NH_DV_ORACLE
    meshes, tips, grasp = _grasp_scene(pkg, n_env, 2500, 11)
    grasp = [g[:k] for g in grasp]
    model = pkg.SimModelDO(meshes, list(range(n_env)), device=gpu, tips=tips, **kw)
    o32 = orc.OracleSim(meshes, list(range(n_env)), np.float32, tips=tips)
    o64 = orc.OracleSim(meshes, list(range(n_env)), np.float64, tips=tips)
    for i, g in enumerate(grasp):
        model.set_grasp(i, g)
        o32.drag[o32.vbase[i] + g] = 1.0
        o64.drag[o64.vbase[i] + g] = 1.0
    rng = np.random.default_rng(5)
    integ = pkg.SimIntegratorDO(gpu)
    cfg = orc.StepConfig(**okw)
    for _ in range(frames):
        delta = rng.normal(0.0, 0.002, size=(n_env, 3)).astype(np.float32)
        dt = torch.from_numpy(delta).to(gpu)
        pkg.lib.ffsrl_move_tools(model.batch, dt.data_ptr(), None, None, model.stream())
        if form == "batched":
            integ.stepModel(model)
        else:
            integ.stepModelPlain(model)
        step = np.repeat(delta, 4, axis=0)
        o32.tips = (o32.tips + step).astype(np.float32)
        o64.tips = o64.tips + step.astype(np.float64)
        o32.step(cfg, 1)
        o64.step(cfg, 1)
    torch.cuda.synchronize()
    assert np.array_equal(model.laparoscopeTip.cpu().numpy().reshape(-1, 3), o32.tips)
    x = model.vertex().cpu().numpy()
    for i in range(n_env):
        s = slice(o32.vbase[i], o32.vbase[i + 1])
        diag = meshes[i].bbox_diagonal()
        e = _report(f"C2 {form} env{i}", x[s], o32.x[s], o64.x[s], diag)
        assert e <= 1e-5 * diag
    # the grasped vertices follow the tool
    g0 = model.vertex()[model.env_slice(0)][torch.as_tensor(grasp[0], device=gpu)].mean(0).cpu().numpy()
    assert np.linalg.norm(g0 - model.laparoscopeTip[0, 2].cpu().numpy()) < 0.1


SETS = {
    "pbd_dv": (dict(distance="pbd", volume="pbd", globalKsDistance=0.7, globalKsVolume=1.0),
               dict(dist_mode=1, dist_k=0.7, vol_mode=1, vol_k=1.0)),
    "xpbd_dv": (dict(distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=1e-7),
                dict(dist_mode=2, vol_mode=2, dist_compliance=1e-6, vol_compliance=1e-7)),
    "nh": (dict(distance=None, volume=None, neoHookean=True, globalVolumeCompliance=2e-5,
                globalDeviatoricCompliance=1e-4), dict(nh_on=1, nh_vol_c=2e-5, nh_dev_c=1e-4)),
    "all": (dict(distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0,
                 neoHookean=True, globalVolumeCompliance=2e-5, globalDeviatoricCompliance=1e-4),
            dict(dist_mode=2, vol_mode=2, dist_compliance=1e-6, vol_compliance=0.0, nh_on=1, nh_vol_c=2e-5,
                 nh_dev_c=1e-4)),
}


@pytest.mark.parametrize("form", ["batched", "plain"])
@pytest.mark.parametrize("iters", [1, 2])
@pytest.mark.parametrize("cset", list(SETS))
def test_single_substep_kernel_parity(pkg, orc, gpu, form, iters, cset):
    """One substep from identical perturbed inputs (positions jittered off rest,
    random velocities, 8 grasped vertices): every constraint family, Jacobi
    averaging, grasp, ground clamp and integrate against the fp32 oracle.  Only
    the dP summation order differs, so the bound is 1e-6 x diag."""
    import dataclasses
    sc = pkg.scenes
    base = [sc.solid("ellipsoid", 3000, seed=21), sc.solid("box", 1500, seed=22)]
    rng = np.random.default_rng(3)
    meshes = [dataclasses.replace(m, verts=(m.verts + rng.uniform(-0.01, 0.01, m.verts.shape)).astype(np.float32))
              for m in base]
    v0 = rng.uniform(-0.05, 0.05, size=(sum(m.num_vertices for m in meshes), 3)).astype(np.float32)
    tips = []
    grasp = []
    for m in meshes:
        tip = sc.top_point(m) + np.array([0.0, 0.02, 0.0], np.float32)
        tips.append(np.repeat(tip[None], 4, axis=0))
        grasp.append(sc.nearest_surface_vertices(m, tip, 8))
    tips = np.stack(tips)
    mk, ok = SETS[cset]
    kw = dict(simSubsteps=1, simFrameRate=600, simConstraintsSteps=iters, environmentGravity=(0.0, -9.81, 0.0),
              environmentVelocityDampening=0.1, globalKsDrag=1.0, **mk)
    model = pkg.SimModelDO(meshes, [0, 1], device=gpu, tips=tips, initialVelocity=v0, **kw)
    o32 = orc.OracleSim(meshes, [0, 1], np.float32, tips=tips, initial_velocity=v0)
    for i, g in enumerate(grasp):
        model.set_grasp(i, g)
        o32.drag[o32.vbase[i] + g] = 1.0
    integ = pkg.SimIntegratorDO(gpu)
    integ.stepModel(model) if form == "batched" else integ.stepModelPlain(model)
    torch.cuda.synchronize()
    o32.step(orc.StepConfig(substeps=1, iterations=iters, fps=600, gravity=(0.0, -9.81, 0.0), damping=0.1,
                            drag_ks=1.0, **ok), 1)
    x = model.vertex().cpu().numpy()
    v = model.v[:, :3].cpu().numpy()
    for i, m in enumerate(meshes):
        s = slice(o32.vbase[i], o32.vbase[i + 1])
        diag = m.bbox_diagonal()
        ex = float(np.abs(x[s] - o32.x[s]).max())
        assert ex <= 1e-6 * diag, (i, ex)
        assert float(np.abs(v[s] - o32.v[s]).max()) <= 600 * 2e-6 * diag
    moved = float(np.abs(o32.x - np.concatenate([m.verts for m in meshes])).max())
    assert moved > 1e-4                                   # the step did something


@pytest.mark.parametrize("block", [1024, 512])
def test_mixed_sizes_alone_equals_in_mix(pkg, gpu, block):
    """Each env's result does not depend on the envs batched with it."""
    sc = pkg.scenes
    specs = [("ellipsoid", 1200), ("box", 5000), ("ellipsoid", 9000), ("box", 2000), ("ellipsoid", 3000)]
    meshes = [sc.solid(s, t, seed=i) for i, (s, t) in enumerate(specs)]
    kw = dict(NH, distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0)
    mix = pkg.SimModelDO(meshes, [4, 0, 3, 1, 2, 0], device=gpu, **kw)
    integ = pkg.SimIntegratorDO(gpu, block=block)
    for _ in range(10):
        integ.stepModel(mix, refit=True)
    for e, k in enumerate([4, 0, 3, 1, 2, 0]):
        alone = pkg.SimModelDO([meshes[k]], [0], device=gpu, **kw)
        for _ in range(10):
            integ.stepModel(alone, refit=True)
        a = alone.vertex().cpu().numpy()
        b = mix.env_vertices(e).cpu().numpy()
        diag = meshes[k].bbox_diagonal()
        err = float(np.abs(a - b).max())
        print(f"\n[alone vs mix] env {e} mesh {k}: {err:.3e} ({err / diag:.2e} diag)")
        # the batched frame sums Jacobi deltas in fixed point (order-free), so an
        # env's result is bit-identical alone and in any mix
        assert np.array_equal(a, b)
        assert torch.equal(alone.bvh_lo, mix.bvh_lo[int(mix.env_nbase[e]):int(mix.env_nbase[e + 1])])


@pytest.mark.parametrize("iters", [1, 2])
def test_constraint_order_does_not_change_results(pkg, gpu, iters):
    """The frame's Jacobi sums are integers, so the visit order of the
    constraints (local, colour-sorted or the mesh's own) gives bit-identical
    positions, velocities and bounds."""
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 4000, seed=11), sc.solid("box", 2500, seed=12)]
    kw = dict(NH, distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0,
              simConstraintsSteps=iters)
    integ = pkg.SimIntegratorDO(gpu)
    out = {}
    top = sc.top_point(meshes[0])
    tips = np.zeros((3, 4, 3), np.float32)
    tips[0, 2] = top + np.array([0.0, 0.05, 0.0])           # lift the grasped vertices 5 cm
    for order in ("local", "color", "mesh"):
        m = pkg.SimModelDO(meshes, [0, 1, 0], device=gpu, constraint_order=order, tips=tips, **kw)
        m.set_grasp(0, sc.nearest_surface_vertices(meshes[0], top, 8))
        for _ in range(8):
            integ.stepModel(m, refit=True)
        out[order] = (m.x.cpu(), m.v.cpu(), m.bvh_lo.cpu(), m.bvh_hi.cpu())
        moved = float((out[order][0] - m.initialVertex.cpu()).abs().max())
        print(f"\n[order {order}] iterations {iters}: max displacement {moved:.3e}")
        assert moved > 1e-4
    for order in ("color", "mesh"):
        for a, b in zip(out["local"], out[order]):
            assert bool(torch.isfinite(a).all())
            assert torch.equal(a, b), order


@pytest.mark.parametrize("collision", [False, True])
def test_block_size_does_not_change_results(pkg, gpu, collision):
    """The 1024-thread loop (indices one stride ahead) and the 256/512-thread
    loop (positions one stride ahead) fuse the same multiply-adds (the
    physics is built with expression-level contraction), so every block size
    gives bit-identical positions, velocities, carried lambdas and bounds;
    with collision, the contact counts too."""
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 3500, seed=41), sc.solid("box", 1500, seed=42)]
    kw = dict(NH, distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0,
              simConstraintsSteps=2)
    top = sc.top_point(meshes[0])
    tips = np.zeros((3, 4, 3), np.float32)
    tips[:, 2] = top + np.array([0.0, 0.05, 0.0])
    if collision:
        # capsule 0 pressed into the top of env 0's and env 2's mesh
        bases = np.zeros((3, 4, 3), np.float32)
        radii = np.zeros((3, 4), np.float32)
        tips[:, 0] = top - np.array([0.0, 0.02, 0.0])
        bases[:, 0] = top + np.array([0.0, 0.3, 0.0])
        radii[:, 0] = 0.05
        kw.update(collision=True, globalKsContact=1.0, capsuleBases=bases, capsuleRadii=radii)
    out = {}
    for block in (1024, 512, 256):
        m = pkg.SimModelDO(meshes, [0, 1, 0], device=gpu, tips=tips, **kw)
        m.set_grasp(0, sc.nearest_surface_vertices(meshes[0], top, 8))
        integ = pkg.SimIntegratorDO(gpu, block=block)
        for _ in range(6):
            integ.stepModel(m, refit=True)
        out[block] = [m.x.cpu(), m.v.cpu(), m.lam_t.cpu(), m.lam_e.cpu(), m.bvh_lo.cpu(), m.bvh_hi.cpu(),
                      m.laparoscopeInCollision.cpu()]
        assert float((out[block][0] - m.initialVertex.cpu()).abs().max()) > 1e-4
    if collision:
        assert int(out[1024][6].sum()) > 0
    for block in (512, 256):
        for k, (a, b) in enumerate(zip(out[1024], out[block])):
            assert bool(torch.isfinite(a.float()).all())
            assert torch.equal(a, b), (block, k)


def test_xpbd_two_iterations_carry_lambda(pkg, orc, gpu):
    mesh, v0 = _c1_model(pkg, 7)
    kw = dict(C1, simConstraintsSteps=2)
    for form in ("batched", "plain"):
        model = pkg.SimModelDO([mesh], [0], device=gpu, initialVelocity=v0, **kw)
        assert model.carries_lambda
        integ = pkg.SimIntegratorDO(gpu)
        for _ in range(30):
            integ.stepModel(model) if form == "batched" else integ.stepModelPlain(model)
        torch.cuda.synchronize()
        cfg = orc.StepConfig(**dict(C1_ORACLE, iterations=2))
        o32 = orc.OracleSim([mesh], [0], np.float32, initial_velocity=v0)
        o64 = orc.OracleSim([mesh], [0], np.float64, initial_velocity=v0)
        o32.step(cfg, 30)
        o64.step(cfg, 30)
        diag = mesh.bbox_diagonal()
        e = _report(f"XPBD I=2 {form}", model.vertex().cpu().numpy(), o32.x, o64.x, diag)
        assert e <= 1e-5 * diag


def test_pbd_reference_do_path(pkg, orc, gpu):
    """The reference DO step's own forms: PBD volume + PBD distance with kS
    (integrator.py:1658-1680) on a negatively oriented mesh (Q4)."""
    mesh = pkg.scenes.cube_grid(7, orientation=-1)
    kw = dict(simSubsteps=8, simFrameRate=30, simConstraintsSteps=1, environmentGravity=(0.0, -9.8, 0.0),
              environmentVelocityDampening=0.1, distance="pbd", volume="pbd", globalKsDistance=1.0,
              globalKsVolume=1.0)
    model = pkg.SimModelDO([mesh], [0], device=gpu, **kw)
    integ = pkg.SimIntegratorDO(gpu)
    for _ in range(40):
        integ.stepModel(model)
    torch.cuda.synchronize()
    cfg = orc.StepConfig(substeps=8, iterations=1, fps=30, gravity=(0.0, -9.8, 0.0), damping=0.1, dist_mode=1,
                         dist_k=1.0, vol_mode=1, vol_k=1.0)
    o32 = orc.OracleSim([mesh], [0], np.float32)
    o64 = orc.OracleSim([mesh], [0], np.float64)
    o32.step(cfg, 40)
    o64.step(cfg, 40)
    diag = mesh.bbox_diagonal()
    e = _report("PBD DO path", model.vertex().cpu().numpy(), o32.x, o64.x, diag)
    assert e <= 1e-5 * diag


def test_reset_restores_initial_state(pkg, gpu):
    mesh, v0 = _c1_model(pkg, 6)
    model = pkg.SimModelDO([mesh, mesh], [0, 1], device=gpu, **C1)
    integ = pkg.SimIntegratorDO(gpu)
    for _ in range(5):
        integ.stepModel(model)
    before = model.env_vertices(1).clone()
    model.resetFixedModel([0])
    torch.cuda.synchronize()
    assert torch.equal(model.env_vertices(0), model.initialVertex[model.env_slice(0), :3])
    assert torch.equal(model.env_vertices(1), before)
    assert not model.v[model.env_slice(0)].any()


@pytest.mark.parametrize("iters", [1, 2])
def test_autotune_keeps_state_and_results(pkg, gpu, tmp_path, iters):
    """Tuning restores the state it timed on, the tuned launch shape gives
    bit-identical frames, and a second tune is served from the cache."""
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 3000, seed=31), sc.solid("box", 1200, seed=32), sc.solid("ellipsoid", 800, seed=33)]
    kw = dict(NH, distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0,
              simConstraintsSteps=iters, lambdaLegacy=True)
    env_mesh = [0, 1, 2, 1, 0, 2, 2, 1]
    ref = pkg.SimModelDO(meshes, env_mesh, device=gpu, **kw)
    tuned = pkg.SimModelDO(meshes, env_mesh, device=gpu, **kw)
    integ_ref = pkg.SimIntegratorDO(gpu)
    integ = pkg.SimIntegratorDO(gpu)
    for _ in range(3):                                  # some state (and carried lambdas) before tuning
        integ_ref.stepModel(ref, refit=True)
        integ.stepModel(tuned, refit=True)
    before = [t.clone() for t in (tuned.x, tuned.v, tuned.bvh_lo)]
    cache = str(tmp_path / "tune.json")
    cfg = integ.autotune(tuned, cache_path=cache)
    print(f"\n[autotune] iterations {iters}: {cfg}")
    assert cfg["block"] in (256, 512, 1024) and cfg["dispatch"] in ("lpt", "table") and cfg["chunk"] in (64, 128, 256)
    for a, b in zip(before, (tuned.x, tuned.v, tuned.bvh_lo)):
        assert torch.equal(a, b)
    for _ in range(4):
        integ_ref.stepModel(ref, refit=True)
        integ.stepModel(tuned, refit=True)
    assert torch.equal(ref.x, tuned.x) and torch.equal(ref.v, tuned.v) and torch.equal(ref.bvh_hi, tuned.bvh_hi)
    assert bool(torch.isfinite(tuned.x).all())
    again = pkg.autotune.Autotuner(cache)
    assert pkg.autotune.key_of(tuned) in again.cache
    assert again.tune(tuned) == cfg


@pytest.mark.parametrize("traversal", [1, 2])
@pytest.mark.parametrize("mode", [0, 1, 2, 3, 5])
def test_rgb_modes_and_normals_bit_exact(pkg, orc, gpu, mode, traversal):
    """Colour renders (grayscale, texture lookup, normal map, vertex colour,
    per-triangle colour, each Lambert-lit by the raw light position) and the
    vertex normals equal the oracle's on the same deformed positions, with
    one ray per lane (1) and with 2x2-ray quads per lane (2)."""
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 3000, seed=41), sc.solid("box", 2000, seed=42)]
    tips = np.stack([np.repeat((sc.top_point(m) + np.array([0.0, 0.05, 0.0]))[None], 4, axis=0) for m in meshes])
    model = pkg.SimModelDO(meshes, [0, 1], device=gpu, tips=tips.astype(np.float32), **NH_DV)
    for i, m in enumerate(meshes):
        model.set_grasp(i, sc.nearest_surface_vertices(m, tips[i, 2], 4))
    integ = pkg.SimIntegratorDO(gpu)
    for _ in range(5):
        integ.stepModel(model, refit=True)
    ren = pkg.RaycastRendererDO(gpu, model, cameraPos=(0.0, 1.6, 5.5), resolution=64, lightPos=(3.0, 2.0, 4.0),
                                lightIntensity=0.3, render_mode=mode, traversal=traversal)
    ren.lookAt((0.0, 0.45, 0.0))
    rgb = ren.render_rgb()
    torch.cuda.synchronize()
    for e, m in enumerate(meshes):
        x = model.env_vertices(e).cpu().numpy()
        nrm = orc.vertex_normals(m.tris, x)
        assert np.array_equal(model.normals[model.env_slice(e), :3].cpu().numpy(), nrm)
        topo = model.topo[e]
        bvh = _bvh_dict(topo)
        mn, mx = orc.refit(bvh, m.tris, x)
        ref = orc.render_rgb(64, 64, mode, np.asarray(ren.camera.origin[:], np.float32), ren.dirs.cpu().numpy(),
                             bvh, mn, mx, m.tris, x, nrm, vcolor=m.colors, tcolor=m.tri_colors,
                             light=np.asarray(ren.camera.light_dir[:], np.float32), intensity=ren.camera.light_intensity,
                             **model.texture_inputs(e))
        got = rgb[e].cpu().numpy()
        lit = float((ref.sum(-1) > 0).mean())
        print(f"\n[rgb mode {mode}] env {e}: lit pixels {lit:.2f}, max |gpu-oracle| {np.abs(got - ref).max():.2e}")
        assert lit > 0.05
        assert np.array_equal(got, ref)
    # the per-env form (one launch per env) gives the same image
    rgb = rgb.clone()
    per_env = ren.render().clone()
    assert torch.equal(per_env, rgb)
    # RGB-D from one traversal: the same colour image and the depth-mode image
    pix, dep = ren.render_rgbd()
    pix, dep = pix.clone(), dep.clone()
    depth = ren.render_depth().clone()
    ren.set_mode(pkg.RenderMode.depth_map)
    assert torch.equal(ren.render()[..., 0], depth)      # depth mode, per env
    torch.cuda.synchronize()
    assert torch.equal(pix, rgb) and torch.equal(dep, depth) and float((depth > 0).float().mean()) > 0.05
