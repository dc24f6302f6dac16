"""The oracle's instrument collision (collideLaparoscopeDO, integrator.py:
981-1050, checkSdfCapsuleCollision :1303-1355), pinned by closed forms: a
sphere-shaped capsule pressing on a triangle moves the triangle out by the
penetration depth, nothing happens out of reach, for a capsule parallel to
the triangle, or to pinned vertices, and the reference's use of the
barycentric coordinates as a point (the selected sphere centre) is kept."""
import numpy as np
import pytest

TRI = np.array([[1.0, 0.0, 2.0], [-0.5, 0.8, 2.0], [-0.5, -0.8, 2.0]])   # z = 2 plane, normal +z
FACE = np.array([[0, 1, 2]])


def _caps(center, axis, radius, half_len):
    """Capsules 0..3 of one env: capsule 0 is the given one, the others far away."""
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    base = np.full((4, 3), 100.0)
    tip = np.full((4, 3), 101.0)
    rad = np.full(4, 0.01)
    base[0] = np.asarray(center) - axis * half_len
    tip[0] = np.asarray(center) + axis * half_len
    rad[0] = radius
    return base, tip, rad


@pytest.mark.parametrize("h", [0.05, 0.15, 0.25])
def test_sphere_capsule_pushes_triangle_out(orc, h):
    r = 0.3
    centroid = TRI.mean(axis=0)
    base, tip, rad = _caps(centroid + [0, 0, h], [0, 0, 1], r, r)      # length 2r: a sphere
    w = np.ones(3)
    out, inc = orc.collide(TRI, FACE, [0], w, base, tip, rad)
    # every vertex moves by (r - h) away from the sphere (down), nothing else changes
    assert np.allclose(out - TRI, [[0, 0, -(r - h)]] * 3, atol=1e-9)
    assert inc.tolist() == [1]
    # stiffness scales the move
    out2, _ = orc.collide(TRI, FACE, [0], w, base, tip, rad, k=0.5)
    assert np.allclose(out2 - TRI, [[0, 0, -0.5 * (r - h)]] * 3, atol=1e-9)


def test_no_response_out_of_reach_parallel_or_pinned(orc):
    centroid = TRI.mean(axis=0)
    w = np.ones(3)
    base, tip, rad = _caps(centroid + [0, 0, 0.31], [0, 0, 1], 0.3, 0.3)   # 1 cm out of reach
    out, inc = orc.collide(TRI, FACE, [0], w, base, tip, rad)
    assert np.array_equal(out, TRI) and inc.tolist() == [0]
    base, tip, rad = _caps(centroid + [0, 0, 0.1], [1, 0, 0], 0.3, 2.0)    # axis parallel to the plane
    out, inc = orc.collide(TRI, FACE, [0], w, base, tip, rad)
    assert np.array_equal(out, TRI) and inc.tolist() == [0]
    base, tip, rad = _caps(centroid + [0, 0, 0.1], [0, 0, 1], 0.3, 0.3)
    out, inc = orc.collide(TRI, FACE, [0], np.array([0.0, 1.0, 1.0]), base, tip, rad)
    assert np.array_equal(out[0], TRI[0]) and not np.array_equal(out[1:], TRI[1:])   # pinned vertex stays
    assert inc.tolist() == [1]


def test_barycentric_coordinates_select_the_sphere_centre(orc):
    """A long capsule along z through the triangle: the reference finds the
    sphere centre closest to the barycentric coordinates of the line/plane
    intersection, a point near the origin, so the centre clamps to the
    capsule's lower end (z = 0.3, 1.7 below the triangle) and there is no
    response, although the capsule crosses the triangle."""
    centroid = TRI.mean(axis=0)
    base = np.full((4, 3), 100.0)
    tip = np.full((4, 3), 101.0)
    rad = np.full(4, 0.01)
    base[0] = [centroid[0], centroid[1], 0.0]
    tip[0] = [centroid[0], centroid[1], 4.0]
    rad[0] = 0.3
    out, inc = orc.collide(TRI, FACE, [0], np.ones(3), base, tip, rad)
    assert np.array_equal(out, TRI) and inc.tolist() == [0]


def test_responses_of_triangles_sum(orc):
    """Two triangles sharing an edge under one sphere: each shared vertex
    receives both triangles' responses (order-free sum)."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
    f = np.array([[0, 1, 2], [1, 3, 2]])
    base, tip, rad = _caps([0.5, 0.5, 0.1], [0, 0, 1], 0.3, 0.3)
    out, inc = orc.collide(v, f, [0, 0], np.ones(4), base, tip, rad)
    assert inc.tolist() == [2]
    d = out - v
    # the contact point is the shared edge's midpoint (barycentrics 0, 1/2, 1/2):
    # each triangle moves the edge's vertices by (r - h) * (1/2) / (1/2), the
    # other vertex not at all; the edge gets both triangles' responses
    r_minus_h = 0.2
    assert np.allclose(d, [[0, 0, 0], [0, 0, -2 * r_minus_h], [0, 0, -2 * r_minus_h], [0, 0, 0]], atol=1e-12)
    single = [orc.collide(v, f[i:i + 1], [0], np.ones(4), base, tip, rad)[0] - v for i in range(2)]
    assert np.allclose(d, single[0] + single[1], atol=1e-12)
