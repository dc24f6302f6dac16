"""The shared-env frame (ffsrl_frame_shared: every CU works on a few
environments, each env's frame cut into chunks taken by the workgroups of
one XCD) against the one-workgroup-per-env frame (ffsrl_frame_envs /
ffsrl_frame): positions, velocities, predicted positions, carried lambdas,
BVH bounds and contact counts equal bit for bit (both sum the Jacobi deltas
as integers), for mixed sizes, XPBD with carried lambdas, the grasp, the
instrument collision, env lists, every block size, and repeated frames (the
chunk counters and sums must be left at zero).  Envs of at most 5,104
vertices take the folded form (one hand-off per unit, the apply folded into
the next pass) unless the model's fold workspace is off; the tests run both
forms, on batches that mix folded and unfolded envs.  The physics itself is
pinned against the oracle by the batched frame's own parity tests.
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


def _meshes(pkg, n, lo, hi, seed):
    sc = pkg.scenes
    specs = sc.mixed_sizes(n, lo, hi, seed=seed)[0]
    return [sc.solid(shape, int(t), seed=seed + i) for i, (shape, t) in enumerate(specs)]


def _pair(pkg, gpu, meshes, kw, grasp=True, **extra):
    sc = pkg.scenes
    tips = np.stack([np.repeat((sc.top_point(m) + np.array([0.0, 0.02, 0.0], np.float32))[None], 4, axis=0)
                     for m in meshes]).astype(np.float32)
    models = []
    for _ in range(2):
        m = pkg.SimModelDO(meshes, list(range(len(meshes))), device=gpu, tips=tips, **kw, **extra)
        if grasp:
            for e in range(1, len(meshes), 2):
                m.set_grasp(e, sc.nearest_surface_vertices(meshes[e], tips[e, 2], 8))
        models.append(m)
    return models


STATE = ("x", "v", "p", "bvh_lo", "bvh_hi")


def _equal(a, b, extra=()):
    """Bit-equal states, and equal status words: a saturated fixed-point
    term (the 8-vertex grasps of the odd envs collapse their tets) is
    reported alike by both forms."""
    torch.cuda.synchronize()
    for k in STATE + tuple(extra) + ("status",):
        ta, tb = getattr(a, k), getattr(b, k)
        assert torch.equal(ta, tb), (k, float((ta - tb).abs().max()))


def _clean(m):
    """Every shared frame completed and left its sums and counters at zero."""
    torch.cuda.synchronize()
    incomplete = m.frame_status() & pkg_status_incomplete()
    assert int(incomplete.abs().sum()) == 0, m.frame_status()
    assert int(m.split_ctr.abs().sum()) == 0 and int(m.split_acc.abs().sum()) == 0
    if getattr(m, "fold_ws", None) is not None:
        assert int(m.fold_ws[:3].abs().sum()) == 0          # the folded form's sum planes


FOLD_LIMIT = 5104


def _folds(meshes):
    """How many of the meshes the folded form takes."""
    return sum(m.num_vertices <= FOLD_LIMIT for m in meshes)


def _clean_sums(m):
    torch.cuda.synchronize()
    assert int(m.split_ctr.abs().sum()) == 0 and int(m.split_acc.abs().sum()) == 0
    if getattr(m, "fold_ws", None) is not None:
        assert int(m.fold_ws[:3].abs().sum()) == 0


def pkg_status_incomplete():
    import importlib
    return importlib.import_module("ff-srl_amd")._lib.STATUS_INCOMPLETE


@pytest.mark.parametrize("block,fold", [(1024, True), (512, True), (256, True), (1024, False), (256, False)])
def test_shared_equals_batched_mixed(pkg, gpu, block, fold):
    meshes = _meshes(pkg, 12, 1000, 30000, seed=5)
    a, b = _pair(pkg, gpu, meshes, NH_DV)
    b.enable_shared(fold=fold)
    ref = pkg.SimIntegratorDO(gpu, block=block)
    sh = pkg.SimIntegratorDO(gpu, block=block, shared=True)
    for f in range(3):
        ref.stepModel(a, refit=True)
        sh.stepModel(b, refit=True)
        _equal(a, b)
        _clean(b)
    assert bool(torch.isfinite(b.x).all())
    print(f"\n[shared] block {block} fold {fold}: 12 envs ({sum(m.num_tets for m in meshes)} tets, "
          f"{_folds(meshes) if fold else 0} folded) equal after 3 frames")


@pytest.mark.parametrize("fold", [True, False])
def test_shared_xpbd_carried_lambdas(pkg, gpu, fold):
    meshes = _meshes(pkg, 6, 2000, 12000, seed=17)
    a, b = _pair(pkg, gpu, meshes, XPBD_I2)
    b.enable_shared(fold=fold)
    assert a.lam_t is not None
    ref, sh = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, shared=True)
    for f in range(3):
        ref.stepModel(a, refit=True)
        sh.stepModel(b, refit=True)
    _equal(a, b, ("lam_t", "lam_e"))
    _clean(b)


@pytest.mark.parametrize("fold", [True, False])
def test_shared_env_list(pkg, gpu, fold):
    meshes = _meshes(pkg, 8, 1000, 15000, seed=23)
    a, b = _pair(pkg, gpu, meshes, NH_DV)
    b.enable_shared(fold=fold)
    envs = torch.tensor([5, 0, 2], dtype=torch.int32, device=gpu)
    x0 = b.x.clone()
    for f in range(2):
        pkg.lib.ffsrl_frame_envs(a.batch, a.params, envs.data_ptr(), 3, 1024, 1, a.stream())
        pkg.lib.ffsrl_frame_shared(b.batch, b.params, envs.data_ptr(), 3, 1024, 1, b.stream())
    _equal(a, b)
    _clean(b)
    # the envs not listed did not move
    for e in (1, 3, 4, 6, 7):
        sl = b.env_slice(e)
        assert torch.equal(b.x[sl], x0[sl])
    # then all envs, and a single env
    pkg.SimIntegratorDO(gpu).stepModel(a, refit=True)
    pkg.SimIntegratorDO(gpu, shared=True).stepModel(b, refit=True)
    one = envs[2:]
    pkg.lib.ffsrl_frame_envs(a.batch, a.params, one.data_ptr(), 1, 1024, 1, a.stream())
    pkg.lib.ffsrl_frame_shared(b.batch, b.params, one.data_ptr(), 1, 1024, 1, b.stream())
    _equal(a, b)
    _clean(b)


def test_shared_collision(pkg, gpu):
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 5000, seed=21 + i) for i in range(3)]
    R = 0.12
    bases, tips = [], []
    for m in meshes:
        q = sc.top_point(m).astype(np.float64)
        c0 = q + np.array([0.0, 0.7 * R, 0.0])
        bases.append(np.stack([c0 - [0, R, 0], q + [-0.15, 0.0, 0.0], q + [0, 3.0, 0], q + [0, 0.5, 0]]))
        tips.append(np.stack([c0 + [0, 0.5 + R, 0], q + [-0.15, 0.6, 0.0], q + [0, 3.5, 0], q + [0, 0.5, 0]]))
    f32 = lambda v: np.asarray(v, np.float32)
    kw = dict(NH_DV, neoHookean=False)
    models = [pkg.SimModelDO(meshes, [0, 1, 2], device=gpu, tips=f32(tips), capsuleBases=f32(bases),
                             capsuleRadii=f32([[R, R, 0.05, 0.0]] * 3), collision=True, globalKsContact=1.0, **kw)
              for _ in range(2)]
    a, b = models
    ref, sh = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, shared=True)
    d = torch.tensor([[0.0, -0.004, 0.0]] * 3, dtype=torch.float32, device=gpu)
    for f in range(4):
        for m in (a, b):
            pkg.lib.ffsrl_move_tools(m.batch, d.data_ptr(), None, None, m.stream())
        ref.stepModel(a, refit=True)
        sh.stepModel(b, refit=True)
    _equal(a, b, ("laparoscopeInCollision",))
    assert int(b.laparoscopeInCollision.sum()) > 0
    _clean(b)


@pytest.mark.parametrize("shared", [False, True])
def test_frame_graph_replays_equal_eager(pkg, gpu, shared):
    """One frame (batched or shared) captured in a HIP graph and replayed four
    times, with eager work and synchronisations between the replays, gives
    after every replay the state of the same number of eager frames (the
    shared frame resets its own chunk counters; nothing in it is a memset)."""
    import ctypes as C
    meshes = _meshes(pkg, 8, 1000, 12000, seed=41)
    a, b = _pair(pkg, gpu, meshes, NH_DV)
    integ = pkg.SimIntegratorDO(gpu, shared=shared)
    lib = pkg.lib
    if shared:
        b.enable_shared()                      # allocation and the XCD probe stay outside the capture
    order = b.dispatch_order()
    torch.cuda.synchronize()
    g = C.c_void_p()
    cs = torch.cuda.Stream()
    with torch.cuda.stream(cs):
        lib.ffsrl_graph_begin(cs.cuda_stream)
        if shared:
            lib.ffsrl_frame_shared(b.batch, b.params, order.data_ptr(), b.numEnvs, 1024, 1, cs.cuda_stream)
        else:
            lib.ffsrl_frame(b.batch, b.params, order.data_ptr(), 1024, 1, cs.cuda_stream)
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
            _equal(a, b)
            if shared:
                _clean(b)
    finally:
        lib.ffsrl_graph_destroy(g)


def test_env_at_the_vertex_limit(pkg, gpu):
    """An env with exactly ffsrl_frame_max_vertices() vertices runs in the
    batched and the shared frame (its LDS and the kernels' static LDS fit in
    160 KB): a 21^3 cube grid padded with pinned, isolated vertices inside its
    own box (same diagonal, so the same fixed-point quantum) evolves exactly
    like the cube alone."""
    sc = pkg.scenes
    cube = sc.cube_grid(21)
    limit = pkg.lib.ffsrl_frame_max_vertices()
    pad = limit - cube.num_vertices
    assert pad > 0
    verts = np.concatenate([cube.verts, np.full((pad, 3), 0.5, np.float32)])
    inv_mass = np.concatenate([cube.inv_mass, np.zeros(pad, np.float32)])
    padded = sc._finish("padded-cube", verts, cube.tets, inv_mass)
    assert padded.num_vertices == limit and np.array_equal(padded.tris, cube.tris)
    kw = dict(NH_DV, neoHookean=False)
    for shared in (False, True):
        ref = pkg.SimModelDO([cube], [0], device=gpu, **kw)
        big = pkg.SimModelDO([padded], [0], device=gpu, **kw)
        integ = pkg.SimIntegratorDO(gpu, shared=shared)
        for _ in range(3):
            integ.stepModel(ref, refit=True)
            integ.stepModel(big, refit=True)
        torch.cuda.synchronize()
        n = cube.num_vertices
        assert torch.equal(big.x[:n], ref.x) and torch.equal(big.v[:n], ref.v), shared
        assert torch.equal(big.bvh_lo, ref.bvh_lo) and torch.equal(big.bvh_hi, ref.bvh_hi), shared
        assert not torch.equal(ref.x, ref.initialVertex)          # the cube moved
        if shared:
            _clean(big)


@pytest.mark.parametrize("fold", [True, False])
def test_abandoned_frame_is_reported_and_cleaned(pkg, gpu, fold):
    """A shared frame whose bounded waits give up: the host sets two envs'
    give-up flags before the launch, so their first wait that does not find
    the pass complete gives up and every workgroup exits normally.  The
    frame is reported (STATUS_INCOMPLETE) for exactly those envs, their
    partial global sums are cleared and every counter is reset; from the
    same starting state the next frame equals the batched frame bit for
    bit, with no status."""
    meshes = _meshes(pkg, 6, 4000, 20000, seed=29)
    a, b = _pair(pkg, gpu, meshes, NH_DV, grasp=False)
    b.enable_shared(fold=fold)
    ref, sh = pkg.SimIntegratorDO(gpu), pkg.SimIntegratorDO(gpu, shared=True)
    sh.stepModel(b, refit=True)
    ref.stepModel(a, refit=True)
    _equal(a, b)
    _clean(b)
    saved = {k: getattr(b, k).clone() for k in STATE}
    b.split_ctr[[1, 4], 2] = 1                        # give-up flags, as a workgroup that gave up leaves them
    sh.stepModel(b, refit=True)
    torch.cuda.synchronize()
    st = b.frame_status().cpu().numpy()
    print(f"\n[shared] status after the abandoned frame: {st.tolist()}")
    assert st[1] & pkg._lib.STATUS_INCOMPLETE and st[4] & pkg._lib.STATUS_INCOMPLETE
    assert (st[[0, 2, 3, 5]] == 0).all()
    _clean_sums(b)
    with pytest.raises(pkg.FFSRLError, match="incomplete shared frame"):
        b.check_status()
    b.clear_status()
    for k, t in saved.items():
        getattr(b, k).copy_(t)
    ref.stepModel(a, refit=True)
    sh.stepModel(b, refit=True)
    _equal(a, b)
    _clean(b)
