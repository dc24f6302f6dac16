"""The oracle's vertex normals and colour-mode shading on the CPU, pinned by
closed forms: symmetric polyhedra (normals known exactly), a flat quad seen
head-on (Lambert term and normal map known), barycentric interpolation
checked against area ratios computed from determinants in float64, and the
texture lookup (texture_interpolation, render.py:124-148) on constant,
linear and numbered textures."""
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
    fc = orc.render_rgb(4, 4, 5, o, dirs, bvh, mn, mx, f, v, n, tcolor=tc, light=light, intensity=inten)
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


# ---------------------------------------------------------------- textures
def test_texture_constant_and_dyadic_lookups_are_exact(orc):
    """A constant texture of 0.5 returns 0.5 exactly wherever the bilinear
    weights are dyadic (side 9: (side - 1) = 8, so coordinates k/16 give
    weights of a few bits and every product and sum is exact)."""
    side = 9
    tex = np.full((side * side, 3), 0.5, np.float32)
    for s in np.arange(0, 17) / 16.0:
        for t in np.arange(0, 17) / 16.0:
            assert np.array_equal(orc.texture_lookup(s, t, tex, 0, side), [0.5, 0.5, 0.5]), (s, t)


def test_texture_reproduces_a_linear_ramp(orc):
    """Bilinear interpolation reproduces an affine function of the texel
    coordinates: with texel[y + x * side] = (x, y, x + 2y) (the reference's
    y + x * width indexing), the lookup at (s, t) is (X, Y, X + 2Y) with
    X = s (side - 1), Y = (1 - t)(side - 1), to float rounding."""
    side = 16
    x, y = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")          # x major: index y + x * side
    tex = np.stack([x, y, x + 2 * y], -1).reshape(-1, 3).astype(np.float32)
    rng = np.random.default_rng(1)
    for s, t in rng.uniform(0.0, 0.999, size=(200, 2)):
        X, Y = s * (side - 1), (1.0 - t) * (side - 1)
        got = orc.texture_lookup(s, t, tex, 0, side, dtype=np.float64)
        assert np.allclose(got, [X, Y, X + 2 * Y], atol=1e-12), (s, t)
        got32 = orc.texture_lookup(s, t, tex, 0, side)
        assert np.allclose(got32, [X, Y, X + 2 * Y], atol=2e-5), (s, t)


def test_texture_index_quirk_and_corners(orc):
    """Texels numbered 0..side^2-1 (value = index): the lookup at a texel
    centre reads texel y + x * side, not y * side + x (the transposed
    reading), with x from s and y from 1 - t; the corners (s, t) = (0, 1)
    and (1, 0) read texels 0 and (side - 1) + (side - 1) * side.  A second
    texture after the first is addressed from its first texel."""
    side = 9
    tex = np.repeat(np.arange(2 * side * side, dtype=np.float32)[:, None], 3, axis=1)
    # s = 0.25 -> x = 2; t = 0.375 -> y = (1 - 0.375) * 8 = 5
    assert orc.texture_lookup(0.25, 0.375, tex, 0, side)[0] == 5 + 2 * side
    assert orc.texture_lookup(0.25, 0.375, tex, 0, side)[0] != 5 * side + 2
    assert orc.texture_lookup(0.0, 1.0, tex, 0, side)[0] == 0
    assert orc.texture_lookup(1.0, 0.0, tex, 0, side)[0] == (side - 1) + (side - 1) * side
    assert orc.texture_lookup(0.25, 0.375, tex, side * side, side)[0] == side * side + 5 + 2 * side
    # half way between x = 2 and 3 at y = 5: the mean of texels 23 and 32
    assert orc.texture_lookup(2.5 / 8, 0.375, tex, 0, side)[0] == 0.5 * (23 + 32)


def test_texture_reads_past_the_table_as_zero(orc):
    """s = 1 puts x0 at the last column and x1 one past it; the reference
    reads that texel with weight 0.  Past the end of the table the oracle
    reads zero, which leaves the result unchanged."""
    side = 4
    tex = np.full((side * side, 3), 0.25, np.float32)
    assert np.array_equal(orc.texture_lookup(1.0, 0.0, tex, 0, side), [0.25, 0.25, 0.25])


def _transposed_lookup(s, t, tex, side):
    """A plausible misreading of render.py:137-141 (row-major y * width + x)."""
    X, Y = s * (side - 1), (1.0 - t) * (side - 1)
    x0, y0 = int(X), int(Y)
    ax, ay = X - x0, Y - y0
    at = lambda x, y: tex[y * side + x] if 0 <= y * side + x < len(tex) else 0.0
    lower = (1 - ax) * at(x0, y0) + ax * at(x0 + 1, y0)
    upper = (1 - ax) * at(x0, y0 + 1) + ax * at(x0 + 1, y0 + 1)
    return (1 - ay) * lower + ay * upper


def test_texture_quirk_test_catches_the_transposed_reading():
    side = 9
    tex = np.arange(side * side, dtype=np.float64)
    assert _transposed_lookup(0.25, 0.375, tex, side) == 5 * side + 2 != 5 + 2 * side


def test_texture_mode_and_fallback_on_a_flat_quad(orc):
    """Mode 1 on the head-on quad: triangle 0 carries texture coordinates and
    a constant texture of 0.5, triangle 1 has no texture and takes the mean
    of its vertex colours (render.py:378-380).  With the light along the
    normal (Lambert term 0.6) the two halves read 0.3 and 0.6 x the mean."""
    v, f, bvh, mn, mx, o, dirs = _quad_scene(orc)
    n = orc.vertex_normals(f, v)
    light, inten = (0.0, 0.0, 2.0), 0.3
    col = np.array([[0.2, 0.4, 0.8], [0.4, 0.4, 0.4], [0.6, 0.1, 0.2], [0.0, 0.0, 0.0]], np.float32)
    tri_uv = np.array([[[0.0, 0.0], [1.0, 0.0], [1.0, 1.0]], [[0.0, 0.0], [1.0, 1.0], [0.0, 1.0]]], np.float32)
    side = 9
    texels = np.full((side * side, 3), 0.5, np.float32)
    out = orc.render_rgb(4, 4, 1, o, dirs, bvh, mn, mx, f, v, n, vcolor=col, light=light, intensity=inten,
                         tri_uv=tri_uv, tri_tex=np.array([0, -1], np.int32), texels=texels,
                         tex_info=np.array([[0, side]], np.int64))
    mean1 = (col[0] + col[2] + col[3]) / 3.0                     # triangle 1 = vertices 0, 2, 3
    fallback = mean1 * 0.6
    px = out.reshape(-1, 3)
    tex_px = np.all(np.abs(px - 0.3) < 1e-6, axis=1)
    fb_px = np.all(np.abs(px - fallback) < 1e-6, axis=1)
    assert tex_px.any() and fb_px.any()
    # every other pixel is a 2x2 mean of rays from both halves: a quarter,
    # half or three quarters of its rays on the textured triangle
    for p in px[~(tex_px | fb_px)]:
        assert any(np.allclose(p, a * 0.3 + (1 - a) * fallback, atol=1e-6) for a in (0.25, 0.5, 0.75)), p
