"""ffsrl_frame_shared_noalloc: the shared-env frame with its global sums and
chunk counters in the library's own device workspace (the call allocates
nothing), against ffsrl_frame_shared with the caller's split_acc /
split_ctr, bit for bit: two calls in a row, the second over fewer envs
listed in another order.  Runs alone in a fresh process and after the other
GPU tests (the workspace keeps the last call's contents, which every frame
leaves at zero)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NH_DV = dict(simSubsteps=10, simFrameRate=60, simConstraintsSteps=2, environmentGravity=(0.0, -9.81, 0.0),
             environmentVelocityDampening=0.1, distance="xpbd", volume="xpbd", distanceCompliance=1e-6,
             volumeComplianceXPBD=0.0, neoHookean=True, globalVolumeCompliance=1.0 / 5e4,
             globalDeviatoricCompliance=1.0 / 1e4, globalKsDrag=1.0)
STATE = ("x", "v", "p", "lam_t", "lam_e", "bvh_lo", "bvh_hi")


def test_noalloc_shared_frame_equals_shared_frame(pkg, gpu):
    sc = pkg.scenes
    specs = sc.mixed_sizes(6, 1000, 12000, seed=3)[0]
    meshes = [sc.solid(shape, int(t), seed=50 + i) for i, (shape, t) in enumerate(specs)]
    tips = np.stack([np.repeat((sc.top_point(m) + np.array([0.0, 0.02, 0.0], np.float32))[None], 4, axis=0)
                     for m in meshes]).astype(np.float32)
    models = [pkg.SimModelDO(meshes, list(range(6)), device=gpu, tips=tips, **NH_DV) for _ in range(2)]
    for m in models:
        for e in (1, 4):
            m.set_grasp(e, sc.nearest_surface_vertices(meshes[e], tips[e, 2], 1))
    ref, noalloc = models
    ref.enable_shared()                                   # the caller's workspace; probes the XCDs
    assert pkg.lib.ffsrl_xcd_count() > 0
    assert getattr(noalloc, "split_acc", None) is None    # the no-allocation frame needs none
    s = torch.cuda.current_stream().cuda_stream
    for envs in ([5, 0, 3, 1, 4, 2], [4, 1, 3]):
        lst = torch.tensor(envs, dtype=torch.int32, device=gpu)
        for _ in range(2):
            pkg.lib.ffsrl_frame_shared(ref.batch, ref.params, lst.data_ptr(), len(envs), 1024, 1, s)
            pkg.lib.ffsrl_frame_shared_noalloc(noalloc.batch, noalloc.params, lst.data_ptr(), len(envs), 1024, 1, s)
        torch.cuda.synchronize()
        for k in STATE:
            a, b = getattr(ref, k), getattr(noalloc, k)
            assert torch.equal(a, b), (envs, k, float((a - b).abs().max()))
        assert int(ref.frame_status().abs().sum()) == 0 and int(noalloc.frame_status().abs().sum()) == 0
        assert bool(torch.isfinite(noalloc.x).all())
        moved = [float((noalloc.env_vertices(e) - noalloc.initialVertex[noalloc.env_slice(e), :3]).abs().max())
                 for e in envs]
        assert min(moved) > 1e-4
        print(f"\n[no-alloc frame] envs {envs}: bit-identical to ffsrl_frame_shared after 2 frames")
    # envs not listed in the second call kept their state in both
    for e in (0, 2, 5):
        assert torch.equal(ref.env_vertices(e), noalloc.env_vertices(e))
