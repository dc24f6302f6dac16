"""The device SAH build (ffsrl_bvh_build_device) against the host build
(ffsrl_bvh_build, itself equal node for node to the NumPy restatement of
bvh.py:228-518 in the oracle): every array equal, for meshes of many sizes
and shapes, degenerate ones included, built in one launch."""
import importlib
import os
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "liver_retraction.npz")


def _meshes(pkg):
    sc = pkg.scenes
    us = importlib.import_module("ff-srl_amd.usd_scene")
    out = [sc.cube_grid(10), sc.cube_grid(3), us.load_scene(FIXTURE).meshes[0]]
    for i, (shape, tets) in enumerate(sc.mixed_sizes(12, 200, 40000, seed=3)[0]):
        out.append(sc.solid(shape, int(tets), seed=100 + i))
    # a single tet (4 faces) and a flat slab whose centroids share one coordinate
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    out.append(sc._finish("tet", v, np.array([[0, 1, 2, 3]], np.int32), np.ones(4)))
    out.append(sc.solid("box", 3000, semi_axes=(1.0, 0.02, 1.0), seed=7))
    return out


def _host(pkg, m):
    return pkg.model.MeshTopology(m)


def test_device_build_equals_host_build(pkg, gpu):
    meshes = _meshes(pkg)
    t0 = time.perf_counter()
    dev = pkg.model.build_bvhs_device(meshes, gpu)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    for m, d in zip(meshes, dev):
        h = _host(pkg, m)
        n = h.num_nodes
        assert d["num_nodes"] == n and d["max_level"] == h.max_level, m.name
        for k in ("left", "first", "count", "level"):
            assert np.array_equal(d[k][:n], getattr(h, k)), (m.name, k)
            unused = {"left": -1, "first": 0, "count": 0, "level": -1}[k]
            assert (d[k][n:] == unused).all(), (m.name, k)
        assert np.array_equal(d["tri_map"], h.tri_map), m.name
    print(f"\n[bvh build] {len(meshes)} meshes ({sum(m.num_tris for m in meshes)} triangles) on the device "
          f"in {t_dev * 1e3:.1f} ms (first call)")


def test_model_with_device_bvh(pkg, gpu):
    sc = pkg.scenes
    meshes = [sc.solid("ellipsoid", 3000, seed=s) for s in range(10)]
    a = pkg.SimModelDO(meshes, device=gpu, bvh_on_device=True)
    b = pkg.SimModelDO(meshes, device=gpu)
    for ta, tb in zip(a.topo, b.topo):
        for k in ("left", "first", "count", "level", "tri_map", "level_nodes", "level_info", "level_off"):
            assert np.array_equal(getattr(ta, k), getattr(tb, k)), k
    integ = pkg.SimIntegratorDO(gpu)
    for m in (a, b):
        integ.stepModel(m, refit=True)
    torch.cuda.synchronize()
    assert torch.equal(a.bvh_lo, b.bvh_lo) and torch.equal(a.bvh_hi, b.bvh_hi)


def test_batched_build_timing(pkg, gpu):
    """Scene admission: 256 meshes (log-uniform 1k-40k tets) built in one
    launch, against the host build of the same meshes on one thread."""
    sc = pkg.scenes
    W = importlib.import_module("ff-srl_amd.workloads")
    specs = [(shape, int(t), 500 + i) for i, (shape, t) in enumerate(sc.mixed_sizes(256, 1000, 40000, seed=9)[0])]
    meshes = W.generate(specs)
    pkg.model.build_bvhs_device(meshes[:4], gpu)            # warm-up (module load)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev = pkg.model.build_bvhs_device(meshes, gpu)
    torch.cuda.synchronize()
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    for m in meshes[:64]:
        pkg.model.build_bvh_host(m)
    t_host = (time.perf_counter() - t0) * len(meshes) / 64
    for m, d in list(zip(meshes, dev))[:8]:
        h = _host(pkg, m)
        assert np.array_equal(d["left"][:h.num_nodes], h.left) and np.array_equal(d["tri_map"], h.tri_map)
    print(f"\n[bvh build] 256 meshes, {sum(m.num_tris for m in meshes)} triangles: device {t_dev * 1e3:.1f} ms "
          f"(incl. transfers), host SAH build (one thread, estimated from 64) {t_host * 1e3:.0f} ms")


def test_simbvh_save_then_load(pkg, gpu, tmp_path):
    """SimBVH(save=True) writes the reference .npz; SimBVH(load=True) on a
    model whose own tree differs (built with 2 bins) installs the saved
    topology on every env of that mesh: same node arrays, and after a frame
    the same bounds and depth image as the model that wrote it."""
    sc = pkg.scenes
    mesh = sc.solid("ellipsoid", 6000, seed=31)
    path = str(tmp_path / "bvh.npz")
    a = pkg.SimModelDO([mesh], [0, 0, 0], device=gpu)
    pkg.SimBVH(gpu, a, save=True, path=path)
    b = pkg.SimModelDO([mesh], [0, 0, 0], device=gpu, bins=2)
    assert not np.array_equal(b.topo[0].left, a.topo[0].left)      # a different tree to start from
    pkg.SimBVH(gpu, b, load=True, path=path)
    for k in ("left", "first", "count", "level", "tri_map", "level_nodes", "level_info", "level_off"):
        assert np.array_equal(getattr(b.topo[0], k), getattr(a.topo[0], k)), k
    integ = pkg.SimIntegratorDO(gpu)
    for m in (a, b):
        for _ in range(3):
            integ.stepModel(m, refit=True)
    torch.cuda.synchronize()
    assert torch.equal(a.x, b.x) and torch.equal(a.bvh_lo, b.bvh_lo) and torch.equal(a.bvh_hi, b.bvh_hi)
    depth = []
    for m in (a, b):
        ren = pkg.RaycastRendererDO(gpu, m, cameraPos=(0.0, 1.6, 5.5), resolution=32)
        ren.lookAt((0.0, 0.3, 0.0))
        depth.append(ren.render_depth().clone())
    assert torch.equal(depth[0], depth[1]) and float((depth[0] > 0).float().mean()) > 0.05
    with pytest.raises(ValueError):
        c = pkg.SimModelDO([sc.solid("box", 2000, seed=3)], [0], device=gpu)
        pkg.SimBVH(gpu, c, load=True, path=path)
