"""The oracle's vertex normals and colour-mode shading on the CPU, pinned by
closed forms: symmetric polyhedra (normals known exactly), a flat quad seen
head-on (Lambert term and normal map known), and barycentric interpolation
checked against area ratios computed from determinants in float64."""
import numpy as np
import pytest


def _octahedron():
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32)
    f = []
    for sx in (0, 1):
        for sy in (2, 3):
            for sz in (4, 5):
                tri = [sx, sy, sz]
                a, b, c = v[tri]
                if np.dot(np.cross(b - a, c - a), a + b + c) < 0:          # outward winding
                    tri = [sx, sz, sy]
                f.append(tri)
    return v, np.array(f, np.int32)


def test_normals_octahedron_exact(orc):
    v, f = _octahedron()
    n = orc.vertex_normals(f, v)
    assert np.allclose(n, v, atol=1e-7)                  # by symmetry the normal is the vertex direction
    n_rev = orc.vertex_normals(f[:, [0, 2, 1]], v)        # reversed winding flips every normal
    assert np.allclose(n_rev, -v, atol=1e-7)


def test_normals_area_weighting_and_isolated_vertex(orc):
    # two triangles sharing vertex 0: one of area 2 in the xy plane, one of area 0.5 in the xz plane
    v = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [1, 0, 0], [0, 0, -1], [5, 5, 5]], np.float32)
    f = np.array([[0, 1, 2], [0, 3, 4]], np.int32)
    n = orc.vertex_normals(f, v, dtype=np.float64)
    s = np.array([0.0, 0.0, 4.0]) + np.cross(v[3] - v[0], v[4] - v[0]).astype(np.float64)   # (0,0,4) + (0,1,0)
    assert np.allclose(n[0], s / np.linalg.norm(s), atol=1e-12)
    assert np.allclose(n[1], [0, 0, 1]) and np.allclose(n[4], [0, 1, 0])
    assert np.array_equal(n[5], [0, 0, 0])               # a vertex of no face stays zero


def _quad_scene(orc, H=4, z_cam=5.0, half=1.0):
    """Unit quad in z = 0 (normal +z), camera on the axis looking down -z,
    rays aimed at an even grid over [-0.75, 0.75]^2 (all hit)."""
    v = np.array([[-half, -half, 0], [half, -half, 0], [half, half, 0], [-half, half, 0]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    bvh = orc.bvh_build(v, f)
    mn, mx = orc.refit(bvh, f, v)
    RW = 2 * H
    xs = (np.arange(RW) + 0.5) / RW * 1.5 - 0.75
    X, Y = np.meshgrid(xs, xs[::-1])
    tgt = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], 1)
    o = np.array([0, 0, z_cam], np.float64)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dirs = np.concatenate([d, np.zeros((len(d), 1))], 1).astype(np.float32)
    return v, f, bvh, mn, mx, o.astype(np.float32), dirs


def test_lambert_and_normal_map_on_a_flat_quad(orc):
    v, f, bvh, mn, mx, o, dirs = _quad_scene(orc)
    n = orc.vertex_normals(f, v)
    assert np.allclose(n, [[0, 0, 1]] * 4)
    light, inten = (0.0, 0.0, 2.0), 0.3                  # raw light position, not normalised (Q17)
    g = orc.render_rgb(4, 4, 0, o, dirs, bvh, mn, mx, f, v, n, light=light, intensity=inten)
    assert np.allclose(g, 0.6, atol=1e-6)                 # 0.3 * dot((0,0,1), (0,0,2))
    nm = orc.render_rgb(4, 4, 2, o, dirs, bvh, mn, mx, f, v, n)
    assert np.allclose(nm, [0.5, 0.5, 1.0], atol=1e-7)
    col = np.array([[0.2, 0.4, 0.8]] * 4, np.float32)
    vc = orc.render_rgb(4, 4, 3, o, dirs, bvh, mn, mx, f, v, n, vcolor=col, light=light, intensity=inten)
    assert np.allclose(vc, col[0] * 0.6, atol=1e-6)
    tc = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]], np.float32)
    fc = orc.render_rgb(4, 4, 1, o, dirs, bvh, mn, mx, f, v, n, tcolor=tc, light=light, intensity=inten)
    # pixels fully inside one triangle take its colour; a light from behind gives black
    r, gch = fc[..., 0], fc[..., 1]
    assert np.allclose(r + gch, 0.6, atol=1e-6) and (r > 0.5).any() and (gch > 0.5).any()
    dark = orc.render_rgb(4, 4, 0, o, dirs, bvh, mn, mx, f, v, n, light=(0.0, 0.0, -1.0), intensity=inten)
    assert np.array_equal(dark, np.zeros_like(dark))
    bright = orc.render_rgb(4, 4, 0, o, dirs, bvh, mn, mx, f, v, n, light=(0.0, 0.0, 10.0), intensity=inten)
    assert np.array_equal(bright, np.ones_like(bright))   # clamped to 1


def test_barycentric_interpolation(orc):
    """A normal (1,0,0) on vertex 0 only: the normal map's red channel is
    0.5 + 0.5 u, with u the barycentric weight of vertex 0 (area ratio)."""
    v, f, bvh, mn, mx, o, dirs = _quad_scene(orc, H=8)
    n = np.zeros((4, 3), np.float32)
    n[0] = (1, 0, 0)
    rgb64 = orc.render_rgb(8, 8, 2, o.astype(np.float64), dirs.astype(np.float64), bvh, mn, mx, f, v, n,
                           dtype=np.float64)
    # expected per ray, then the 2x2 box
    RW = 16
    ray = np.zeros((RW * RW,))
    o64 = o.astype(np.float64)
    for r in range(RW * RW):
        d = dirs[r, :3].astype(np.float64)
        t = -o64[2] / d[2]
        p = o64 + t * d
        for tri in f:
            A, B, Cc = v[tri].astype(np.float64)
            area = np.linalg.det(np.stack([B - A, Cc - A, [0, 0, 1]]))
            w0 = np.linalg.det(np.stack([B - p, Cc - p, [0, 0, 1]])) / area    # weight of the triangle's first vertex
            w1 = np.linalg.det(np.stack([Cc - p, A - p, [0, 0, 1]])) / area
            w2 = 1 - w0 - w1
            if min(w0, w1, w2) >= -1e-12:
                ray[r] = 0.5 + 0.5 * (w0 if tri[0] == 0 else 0.0)
                break
    exp = ray.reshape(RW, RW).reshape(8, 2, 8, 2).mean(axis=(1, 3))
    assert np.allclose(rgb64[..., 0], exp, atol=1e-9)
    assert np.allclose(rgb64[..., 1], 0.5, atol=1e-12) and np.allclose(rgb64[..., 2], 0.5, atol=1e-12)
