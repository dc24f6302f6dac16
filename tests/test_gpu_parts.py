"""The partitioned frame (ffsrl_frame_parts: each env's constraints cut into
parts, one workgroup per part for the whole frame, the parts' state in LDS
and only shared vertices' sums exchanged through global planes) against the
one-workgroup-per-env frame: positions, velocities, predicted positions,
carried lambdas, BVH bounds and status equal bit for bit (both sum the
Jacobi deltas as integers), for mixed sizes with the grasp, XPBD with
carried lambdas, PBD and edge-only schemes, env lists, parts small enough to
share a third of their vertices, and repeated frames (the planes and
counters must be left at zero).  An env above the batched frame's vertex
limit, which only this frame and the plain form can run, is checked against
the float32 oracle.  A frame whose waits give up is reported and cleaned.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NH_DV = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=1, environmentGravity=(0.0, -9.81, 0.0),
             environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
             volumeComplianceXPBD=0.0, neoHookean=True, globalVolumeCompliance=1.0 / 5e4,
             globalDeviatoricCompliance=1.0 / 1e4, globalKsDrag=1.0)
XPBD_I2 = dict(NH_DV, neoHookean=False, simConstraintsSteps=2, simSubsteps=5)
PBD_DV = dict(NH_DV, distance="pbd", volume="pbd", neoHookean=False, simSubsteps=8)
PBD_EDGES = dict(PBD_DV, volume=None)
NH_ONLY = dict(NH_DV, distance=None, volume=None)
STATE = ("x", "v", "p", "bvh_lo", "bvh_hi", "status")


def _meshes(pkg, n, lo, hi, seed):
    sc = pkg.scenes
    specs = sc.mixed_sizes(n, lo, hi, seed=seed)[0]
    return [sc.solid(shape, int(t), seed=seed + i) for i, (shape, t) in enumerate(specs)]


def _pair(pkg, gpu, meshes, kw, grasp=8):
    sc = pkg.scenes
    tips = np.stack([np.repeat((sc.top_point(m) + np.array([0.0, 0.02, 0.0], np.float32))[None], 4, axis=0)
                     for m in meshes]).astype(np.float32)
    models = []
    for _ in range(2):
        m = pkg.SimModelDO(meshes, list(range(len(meshes))), device=gpu, tips=tips, **kw)
        if grasp:
            for e in range(1, len(meshes), 2):
                m.set_grasp(e, sc.nearest_surface_vertices(meshes[e], tips[e, 2], grasp))
        models.append(m)
    return models


def _equal(a, b, extra=()):
    torch.cuda.synchronize()
    for k in STATE + tuple(extra):
        ta, tb = getattr(a, k), getattr(b, k)
        assert torch.equal(ta, tb), (k, float((ta.float() - tb.float()).abs().max()))


def _clean(pkg, m, complete=True):
    """The frame left the planes and counters at zero (and completed)."""
    torch.cuda.synchronize()
    if complete:
        assert int((m.frame_status() & pkg._lib.STATUS_INCOMPLETE).abs().sum()) == 0, m.frame_status()
    assert int(m.p_ctr[:4 * m.numEnvs + 17].abs().sum()) == 0 and int(m.p_sums.abs().sum()) == 0


def _shared_fraction(m):
    v = m.p_verts.cpu().numpy().view(np.uint32)
    return float(((v & 0x40000000) != 0).mean())


@pytest.mark.parametrize("tets_per_part", [1500, 5120, 1 << 30])
def test_parts_equal_batched_mixed(pkg, gpu, tets_per_part):
    """1 << 30 tets per part: every env of at most 5,088 vertices in one part
    (the positions-in-LDS form of the one-workgroup frame), larger envs in as
    few parts as LDS allows."""
    meshes = _meshes(pkg, 12, 1000, 30000, seed=5)
    a, b = _pair(pkg, gpu, meshes, NH_DV)
    b.enable_parts(tets_per_part=tets_per_part)
    ref, pt = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, parts=True)
    for f in range(3):
        ref.stepModel(a, refit=True)
        pt.stepModel(b, refit=True)
        _equal(a, b)
        _clean(pkg, b)
    assert bool(torch.isfinite(b.x).all())
    print(f"\n[parts] {tets_per_part} tets per part: parts per mesh {b.parts_per_mesh}, "
          f"{100 * _shared_fraction(b):.0f}% shared vertex entries; equal after 3 frames")


@pytest.mark.parametrize("kw,name", [(XPBD_I2, "xpbd-i2"), (PBD_DV, "pbd"), (PBD_EDGES, "pbd-edges"),
                                     (NH_ONLY, "nh")])
def test_parts_schemes(pkg, gpu, kw, name):
    meshes = _meshes(pkg, 6, 2000, 12000, seed=17)
    a, b = _pair(pkg, gpu, meshes, kw, grasp=1)
    b.enable_parts(tets_per_part=2000)
    ref, pt = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, parts=True)
    extra = ("lam_t", "lam_e") if a.lam_t is not None else ()
    for f in range(3):
        ref.stepModel(a, refit=True)
        pt.stepModel(b, refit=True)
    _equal(a, b, extra)
    _clean(pkg, b)
    assert not torch.equal(b.x, b.initialVertex)


def test_parts_env_list(pkg, gpu):
    meshes = _meshes(pkg, 8, 1000, 15000, seed=23)
    a, b = _pair(pkg, gpu, meshes, NH_DV)
    b.enable_parts(tets_per_part=3000)
    envs = torch.tensor([5, 0, 2], dtype=torch.int32, device=gpu)
    x0 = b.x.clone()
    for f in range(2):
        pkg.lib.ffsrl_frame_envs(a.batch, a.params, envs.data_ptr(), 3, 1024, 1, a.stream())
        pkg.lib.ffsrl_frame_parts(b.batch, b.params, b.parts, envs.data_ptr(), 3, 1, b.stream())
    _equal(a, b)
    _clean(pkg, b)
    for e in (1, 3, 4, 6, 7):
        s = b.env_slice(e)
        assert torch.equal(b.x[s], x0[s])                   # envs not listed are untouched


def test_parts_follow_a_new_constraint_order(pkg, gpu):
    """The partition is rebuilt when the constraint order changes (the
    autotuner's chunk knob); the frames stay equal to the batched frame."""
    meshes = _meshes(pkg, 4, 3000, 9000, seed=31)
    a, b = _pair(pkg, gpu, meshes, NH_DV)
    b.enable_parts(tets_per_part=2500)
    old = b.p_tets16.clone()
    for m in (a, b):
        m.set_constraint_order("local", chunk=256)
    assert not torch.equal(old, b.p_tets16)
    ref, pt = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, parts=True)
    for f in range(2):
        ref.stepModel(a, refit=True)
        pt.stepModel(b, refit=True)
    _equal(a, b)
    _clean(pkg, b)


def test_parts_env_above_the_vertex_limit(pkg, orc, gpu):
    """A 64k-tet ellipsoid (above the batched frame's 10,208 vertices, so
    only the plain form ran it before) in 13 parts: positions within
    1e-5 x diag of the float32 oracle after 2 frames (the fp32-fp64 gap is
    printed), BVH bounds equal the oracle's refit of the same positions."""
    sc = pkg.scenes
    mesh = sc.solid("ellipsoid", 64000, seed=3)
    assert mesh.num_vertices > pkg.lib.ffsrl_frame_max_vertices()
    tip = (sc.top_point(mesh) + np.array([0.0, 0.02, 0.0], np.float32)).astype(np.float32)
    tips = np.repeat(tip[None], 4, axis=0)[None].astype(np.float32)
    kw = dict(NH_DV, neoHookean=False, volume="xpbd", distance="xpbd")
    m = pkg.SimModelDO([mesh], [0], device=gpu, tips=tips, **kw)
    g = sc.nearest_surface_vertices(mesh, tip, 1)
    m.set_grasp(0, g)
    with pytest.raises(pkg.FFSRLError, match="LDS capacity"):
        pkg.SimIntegratorDO(gpu).stepModel(m)
    m.enable_parts()
    print(f"\n[parts] {mesh.num_tets} tets, {mesh.num_vertices} vertices: {m.parts_per_mesh[0]} parts, "
          f"largest {m.parts.max_local} local vertices")
    assert m.parts.max_solo == 0
    cfg = dict(substeps=10, iterations=1, fps=60, gravity=(0.0, -9.81, 0.0), damping=0.1, dist_mode=2,
               vol_mode=2, nh_on=0, nh_vol_c=1.0 / 5e4, nh_dev_c=1.0 / 1e4, dist_compliance=1e-6,
               vol_compliance=0.0, drag_ks=1.0)
    o32 = orc.OracleSim([mesh], [0], np.float32, tips=tips)
    o64 = orc.OracleSim([mesh], [0], np.float64, tips=tips)
    for o in (o32, o64):
        o.drag[np.asarray(g)] = 1.0
    integ = pkg.SimIntegratorDO(gpu, parts=True)
    for f in range(2):
        integ.stepModel(m, refit=True)
        o32.step(orc.StepConfig(**cfg), 1)
        o64.step(orc.StepConfig(**cfg), 1)
    _clean(pkg, m)
    x = m.vertex().cpu().numpy()
    diag = mesh.bbox_diagonal()
    err, gap = float(np.abs(x - o32.x).max()) / diag, float(np.abs(o32.x - o64.x).max()) / diag
    print(f"[parts] |gpu-o32| {err:.2e} diag, |o32-o64| {gap:.2e} diag")
    assert err <= 1e-5
    t = m.topo[0]
    mn, mx = orc.refit({"left": t.left, "first": t.first, "count": t.count, "level": t.level,
                        "tri_map": t.tri_map}, mesh.tris, x)
    assert np.array_equal(m.bvh_lo[:, :3].cpu().numpy(), mn) and np.array_equal(m.bvh_hi[:, :3].cpu().numpy(), mx)


def test_parts_abandoned_frame_is_reported_and_cleaned(pkg, gpu):
    """A partitioned frame whose bounded waits give up: the host sets two
    envs' give-up flags before the launch, so a part that does not find its
    env's pass complete gives up and every workgroup exits normally.  The
    frame is reported (STATUS_INCOMPLETE) for exactly those envs, their
    planes are cleared and every counter is reset; from the same starting
    state the next frame equals the batched frame bit for bit."""
    meshes = _meshes(pkg, 6, 4000, 20000, seed=29)
    a, b = _pair(pkg, gpu, meshes, NH_DV, grasp=0)
    b.enable_parts(tets_per_part=2000)
    ref, pt = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, parts=True)
    ref.stepModel(a, refit=True)
    pt.stepModel(b, refit=True)
    _equal(a, b)
    _clean(pkg, b)
    saved = {k: getattr(b, k).clone() for k in STATE if k != "status"}
    b.p_ctr[4 * 1 + 2] = 1
    b.p_ctr[4 * 4 + 2] = 1
    pt.stepModel(b, refit=True)
    torch.cuda.synchronize()
    st = b.frame_status().cpu().numpy()
    print(f"\n[parts] status after the abandoned frame: {st.tolist()}")
    assert st[1] & pkg._lib.STATUS_INCOMPLETE and st[4] & pkg._lib.STATUS_INCOMPLETE
    assert (st[[0, 2, 3, 5]] == 0).all()
    _clean(pkg, b, complete=False)
    b.clear_status()
    for k, t in saved.items():
        getattr(b, k).copy_(t)
    ref.stepModel(a, refit=True)
    pt.stepModel(b, refit=True)
    _equal(a, b)
    _clean(pkg, b)


@pytest.mark.parametrize("collision", [False, True])
def test_parts_graph_replays_equal_eager(pkg, gpu, collision):
    """The partitioned frame (claim table and frame) captured in a HIP graph
    and replayed four times, with eager work and synchronisations between
    the replays, gives after every replay the state of the same number of
    eager batched frames; the frame resets its own counters and planes (no
    memset), with the collision pass's second exchange per pass too."""
    import ctypes as C
    lib = pkg.lib
    if collision:
        sc = pkg.scenes
        meshes = [sc.solid("ellipsoid", t, seed=71 + i) for i, t in enumerate((5000, 3000, 12000, 1500))]
        a, b = _collision_models(pkg, gpu, meshes, NH_DV)
    else:
        meshes = _meshes(pkg, 8, 1000, 12000, seed=41)
        a, b = _pair(pkg, gpu, meshes, NH_DV)
    b.enable_parts(tets_per_part=3000)          # partition, allocation and the XCD probe outside the capture
    assert max(b.parts_per_mesh) > 1 and min(b.parts_per_mesh) == 1
    integ = pkg.SimIntegratorDO(gpu)
    order = b.dispatch_order()
    torch.cuda.synchronize()
    g = C.c_void_p()
    cs = torch.cuda.Stream()
    with torch.cuda.stream(cs):
        lib.ffsrl_graph_begin(cs.cuda_stream)
        lib.ffsrl_frame_parts(b.batch, b.params, b.parts, order.data_ptr(), b.numEnvs, 1, cs.cuda_stream)
        lib.ffsrl_graph_end(cs.cuda_stream, C.byref(g))
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x)                   # capturing ran nothing
    try:
        for k in range(4):
            integ.stepModel(a, refit=True)
            with torch.cuda.stream(cs):
                lib.ffsrl_graph_launch(g, cs.cuda_stream)
            torch.cuda.synchronize()
            _ = float(b.x[:, :3].sum())                # eager work between the replays
            _equal(a, b, ("laparoscopeInCollision",) if collision else ())
            _clean(pkg, b)
    finally:
        lib.ffsrl_graph_destroy(g)
    if collision:
        assert int(b.laparoscopeInCollision.sum()) > 0


def _collision_models(pkg, gpu, meshes, kw, R=0.12):
    """Two models of the same envs: capsule 0 pressing on each env's highest
    surface vertex, capsule 1 beside it, capsule 2 far above."""
    sc = pkg.scenes
    bases, tips = [], []
    for m in meshes:
        q = sc.top_point(m).astype(np.float64)
        c0 = q + np.array([0.0, 0.7 * R, 0.0])
        bases.append(np.stack([c0 - [0, R, 0], q + [-0.15, 0.0, 0.0], q + [0, 3.0, 0], q + [0, 0.5, 0]]))
        tips.append(np.stack([c0 + [0, 0.5 + R, 0], q + [-0.15, 0.6, 0.0], q + [0, 3.5, 0], q + [0, 0.5, 0]]))
    f32 = lambda v: np.asarray(v, np.float32)
    n = len(meshes)
    return [pkg.SimModelDO(meshes, list(range(n)), device=gpu, tips=f32(tips), capsuleBases=f32(bases),
                           capsuleRadii=f32([[R, R, 0.05, 0.0]] * n), collision=True, globalKsContact=1.0, **kw)
            for _ in range(2)]


@pytest.mark.parametrize("tets_per_part,kw,name", [(600, NH_DV, "nh-600"), (1500, XPBD_I2, "xpbd-i2-1500"),
                                                   (1 << 30, NH_DV, "nh-whole")])
def test_parts_collision_equal_batched(pkg, gpu, tets_per_part, kw, name):
    """With the instrument collision pass (capsules pressing 4 mm per frame
    into ellipsoids): state, carried lambdas and the per-env contact counts
    equal the batched frame's bit for bit.  Each part tests its own surface
    triangles; sums of vertices shared between parts cross a second exchange
    per pass.  1 << 30 tets per part: every env in one part."""
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", t, seed=21 + i) for i, t in enumerate((5000, 3000, 12000))]
    a, b = _collision_models(pkg, gpu, meshes, kw)
    b.enable_parts(tets_per_part=tets_per_part)
    # surface triangles with a vertex held by several parts
    v = b.p_verts.cpu().numpy().view(np.uint32)
    rows = b.p_rows.cpu().numpy()
    t16 = b.p_tris16.cpu().numpy()
    crossing = 0
    for r in rows:
        w = t16[r[6]:r[6] + r[7]].view(np.int64)         # 3 x 16 bits: the high bits are zero
        ids = np.stack([(w >> (16 * k)) & 0xffff for k in range(3)])
        assert (ids < r[5]).all()                          # local ids inside the part's vertex list
        crossing += int(((v[r[4] + ids] & 0x40000000) != 0).any(axis=0).sum())
    if tets_per_part < (1 << 30):
        assert crossing > 0
    else:
        assert max(b.parts_per_mesh) == 1
    ref, pt = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, parts=True)
    extra = ("laparoscopeInCollision",) + (("lam_t", "lam_e") if a.lam_t is not None else ())
    d = torch.tensor([[0.0, -0.004, 0.0]] * len(meshes), dtype=torch.float32, device=gpu)
    for f in range(4):
        for m in (a, b):
            pkg.lib.ffsrl_move_tools(m.batch, d.data_ptr(), None, None, m.stream())
        ref.stepModel(a, refit=True)
        pt.stepModel(b, refit=True)
        _equal(a, b, extra)
        _clean(pkg, b)
    contacts = b.laparoscopeInCollision.cpu().numpy()
    print(f"\n[parts collision {name}] parts per mesh {b.parts_per_mesh}, {crossing} triangles with a shared "
          f"vertex, contacts {contacts.tolist()}")
    assert (contacts > 0).all()


def test_autotune_with_collision(pkg, gpu, tmp_path):
    """The autotuner times the partitioned frame with the collision pass too
    (its trials are finite), restores the state and contact counts it timed
    on, and its choice gives the batched frame's results bit for bit."""
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", t, seed=61 + i) for i, t in enumerate((5000, 3000, 12000))]
    a, b = _collision_models(pkg, gpu, meshes, NH_DV)
    ref, integ = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu)
    d = torch.tensor([[0.0, -0.004, 0.0]] * len(meshes), dtype=torch.float32, device=gpu)
    for _ in range(2):
        for m in (a, b):
            pkg.lib.ffsrl_move_tools(m.batch, d.data_ptr(), None, None, m.stream())
        ref.stepModel(a, refit=True)
        integ.stepModel(b, refit=True)
    before = [t.clone() for t in (b.x, b.v, b.laparoscopeInCollision)]
    cache = str(tmp_path / "tune.json")
    cfg = integ.autotune(b, cache_path=cache)
    trials = pkg.autotune.Autotuner(cache).cache[pkg.autotune.key_of(b)]["trials"]
    parted = [t for t in trials if t["parts"]]
    print(f"\n[autotune collision] {cfg}; {len(parted)} partitioned trials, best "
          f"{min(t['ms'] for t in parted):.3f} ms against {min(t['ms'] for t in trials if not t['parts']):.3f} ms")
    assert parted and all(np.isfinite(t["ms"]) for t in parted)
    for x, y in zip(before, (b.x, b.v, b.laparoscopeInCollision)):
        assert torch.equal(x, y)
    for _ in range(3):
        for m in (a, b):
            pkg.lib.ffsrl_move_tools(m.batch, d.data_ptr(), None, None, m.stream())
        ref.stepModel(a, refit=True)
        integ.stepModel(b, refit=True)
    _equal(a, b, ("laparoscopeInCollision",))
    if integ.parts:
        _clean(pkg, b)
    assert int(b.laparoscopeInCollision.sum()) > 0
