"""Runs the host-side C and C++ of the library or of the oracle on small
meshes and prints one SHA-1 digest of every output.

    python tests/_host_san_driver.py product   # the library FFSRL_LIB names
    python tests/_host_san_driver.py oracle    # the oracle FFSRL_ORACLE_LIB names

tests/test_host_sanitizers.py runs it on the normal builds and on the
AddressSanitizer + UndefinedBehaviorSanitizer builds (with the sanitizer
runtime preloaded) and compares the digests.  It imports numpy and loads the
modules it needs by file path, so neither torch nor a GPU is involved; the
library calls made here run on the host only.
"""
import ctypes as C
import hashlib
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


scenes = _load("ffsrl_scenes", os.path.join(ROOT, "ff-srl_amd", "scenes.py"))
H = hashlib.sha1()


def put(*arrays):
    for a in arrays:
        H.update(np.ascontiguousarray(a).tobytes())


def meshes():
    rng = np.random.default_rng(5)
    soup_v = rng.random((40, 3)).astype(np.float32)
    out = [(m.verts, m.tris, m.tets, m.edges) for m in
           (scenes.cube_grid(4), scenes.solid("ellipsoid", 2000, seed=1), scenes.solid("box", 1500, seed=3))]
    out.append((soup_v, rng.integers(0, 40, (60, 3)).astype(np.int32), None, None))       # random soup
    same = np.repeat(np.array([[0, 1, 2]], np.int32), 9, axis=0)                          # equal centroids
    out.append((soup_v, same, None, None))
    out.append((soup_v, np.array([[3, 4, 5]], np.int32), None, None))                     # one triangle
    return out


def product():
    lib = C.CDLL(os.environ["FFSRL_LIB"])
    P = lambda a: C.c_void_p(a.ctypes.data)
    i32 = lambda a: np.ascontiguousarray(a, np.int32)
    lib.ffsrl_last_error.restype = C.c_char_p
    for verts, tris, tets, edges in meshes():
        verts = np.ascontiguousarray(verts, np.float32)
        tris = i32(tris)
        nv, nf = verts.shape[0], tris.shape[0]
        for bins in (2, 8):
            arrs = [np.zeros(2 * nf, np.int32) for _ in range(4)]
            tri_map = np.zeros(nf, np.int32)
            nn, ml = C.c_int32(), C.c_int32()
            r = lib.ffsrl_bvh_build(P(verts), C.c_int32(nv), P(tris), C.c_int32(nf), C.c_int32(bins),
                                    *[P(a) for a in arrs], P(tri_map), C.byref(nn), C.byref(ml))
            assert r == 0, lib.ffsrl_last_error()
            put(np.int32([nn.value, ml.value]), *[a[:nn.value] for a in arrs], tri_map)
        for elems, arity in ((tets, 4), (edges, 2), (tris, 3)):
            if elems is None:
                continue
            elems = i32(elems)
            n = elems.shape[0]
            color, ncol = np.zeros(n, np.int32), C.c_int32()
            r = lib.ffsrl_color_constraints(P(elems), C.c_int64(n), C.c_int32(arity), C.c_int32(nv), P(color),
                                            C.byref(ncol))
            assert r == 0, lib.ffsrl_last_error()
            put(color, np.int32([ncol.value]))
            for chunk in (64, 128, 256):
                perm = np.zeros(n, np.int64)
                r = lib.ffsrl_order_constraints(P(elems), C.c_int64(n), C.c_int32(arity), P(verts), C.c_int32(nv),
                                                C.c_int32(chunk), P(perm))
                assert r == 0, lib.ffsrl_last_error()
                assert np.array_equal(np.sort(perm), np.arange(n))
                put(perm)
    # argument checks and the error message path
    bad = i32([[0, 1, 99]])
    v3 = np.zeros((3, 3), np.float32)
    z = np.zeros(4, np.int32)
    nn, ml = C.c_int32(), C.c_int32()
    assert lib.ffsrl_bvh_build(P(v3), 3, P(bad), 1, 8, P(z), P(z), P(z), P(z), P(z), C.byref(nn), C.byref(ml)) != 0
    assert b"out of range" in lib.ffsrl_last_error()
    assert lib.ffsrl_order_constraints(P(bad), C.c_int64(1), C.c_int32(9), P(v3), 3, 64, P(z)) != 0
    assert b"order_constraints" in lib.ffsrl_last_error()


def oracle():
    orc = _load("ffsrl_oracle", os.path.join(ROOT, "oracle", "oracle.py"))
    mesh = scenes.solid("ellipsoid", 600, seed=2)
    tip = scenes.top_point(mesh) + np.float32([0.0, 0.02, 0.0])
    tips = np.repeat(tip[None], 4, axis=0).astype(np.float32)
    grasp = scenes.nearest_surface_vertices(mesh, tip, 2)
    bases = tips + np.float32([[0.0, 0.5, 0.0], [-0.03, 0.08, 0.0], [0.03, 0.08, 0.0], [0.0, 0.0, 0.0]])
    radii = np.float32([0.03, 0.015, 0.015, 0.0])
    cfgs = [orc.StepConfig(substeps=4, iterations=2, dist_mode=2, vol_mode=2, dist_compliance=1e-6,
                           vol_compliance=0.0, nh_on=1, nh_vol_c=1 / 5e4, nh_dev_c=1 / 1e4, collide=1),
            orc.StepConfig(substeps=3, iterations=1, dist_mode=1, dist_k=1.0, vol_mode=1, vol_k=1.0)]
    for dt in (np.float32, np.float64):
        for cfg in cfgs:
            s = orc.OracleSim([mesh, mesh], [0, 0], dt, tips=np.concatenate([tips, tips]),
                              caps_base=np.concatenate([bases, bases]), caps_radius=np.concatenate([radii, radii]))
            s.drag[np.asarray(grasp)] = 1.0
            for f in range(3):
                s.tips = (s.tips + np.float32([0.0, 0.01, 0.0])).astype(dt)
                s.step(cfg, 1)
            put(s.x, s.v, s.lam_e, s.lam_t, s.in_collision)
    # BVH build (NumPy restatement), refit, normals and every render mode
    v = s.x[:mesh.num_vertices].astype(np.float32)
    bvh = orc.bvh_build(v, mesh.tris, bins=8)
    mn, mx = orc.refit(bvh, mesh.tris, v)
    put(mn, mx)
    H_, W_ = 8, 8
    g = np.stack(np.meshgrid(np.linspace(-0.4, 0.4, 2 * W_), np.linspace(0.4, -0.4, 2 * H_)), -1).reshape(-1, 2)
    d = np.concatenate([g, -np.ones((g.shape[0], 1))], axis=1)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    origin = np.float32([0.0, 0.3, 3.0])
    depth, hit, t, _ = orc.render_depth(H_, W_, origin, d, bvh, mn, mx, mesh.tris, v)
    put(depth, hit, t)
    n = orc.vertex_normals(mesh.tris, v)
    rng = np.random.default_rng(3)
    vcol = rng.random((mesh.num_vertices, 3)).astype(np.float32)
    tcol = rng.random((mesh.num_tris, 3)).astype(np.float32)
    uv = rng.random((mesh.num_tris, 3, 2)).astype(np.float32)
    ttex = np.where(np.arange(mesh.num_tris) % 3 == 0, -1, 0).astype(np.int32)
    texels = rng.random((16 * 16, 3)).astype(np.float32)
    for mode in (0, 1, 2, 3, 5):
        put(orc.render_rgb(H_, W_, mode, origin, d, bvh, mn, mx, mesh.tris, v, n, vcolor=vcol, tcolor=tcol,
                           tri_uv=uv, tri_tex=ttex, texels=texels, tex_info=[[0, 16]]))
    put(np.float32([orc.texture_lookup(s_, t_, texels, 0, 16)[k] for s_ in (0.0, 0.37, 0.99)
                    for t_ in (0.01, 0.5, 1.0) for k in range(3)]))


def short_buffer():
    """A planted mistake: the triangle map one element short.  The builder
    writes its last entry past the end; under ASan that is a reported
    heap-buffer-overflow (the array is large enough to come from malloc)."""
    lib = C.CDLL(os.environ["FFSRL_LIB"])
    P = lambda a: C.c_void_p(a.ctypes.data)
    m = scenes.solid("ellipsoid", 2000, seed=1)
    verts, tris = np.ascontiguousarray(m.verts, np.float32), np.ascontiguousarray(m.tris, np.int32)
    nf = tris.shape[0]
    arrs = [np.zeros(2 * nf, np.int32) for _ in range(4)]
    tri_map = np.zeros(nf - 1, np.int32)
    nn, ml = C.c_int32(), C.c_int32()
    lib.ffsrl_bvh_build(P(verts), C.c_int32(verts.shape[0]), P(tris), C.c_int32(nf), C.c_int32(8),
                        *[P(a) for a in arrs], P(tri_map), C.byref(nn), C.byref(ml))


if __name__ == "__main__":
    {"product": product, "oracle": oracle, "short_buffer": short_buffer}[sys.argv[1]]()
    print("digest", H.hexdigest())
