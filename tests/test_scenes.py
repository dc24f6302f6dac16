"""Scene generator and mesh preparation (host side)."""
import ctypes as C

import numpy as np
import pytest


def test_cube_counts_and_conforming(pkg):
    m = pkg.scenes.cube_grid(10)
    # 9^3 cells x 5 tets; unique tet edges; 6 faces x 81 quads x 2 boundary tris
    assert (m.num_vertices, m.num_tets, m.num_edges, m.num_tris) == (1000, 3645, 5130, 972)
    assert int((m.inv_mass == 0).sum()) == 100                      # bottom face pinned
    assert abs(float(m.tet_rest_volume.astype(np.float64).sum()) - 1.0) < 1e-5   # tets tile the cube
    # every interior face is shared by exactly two tets (conforming split)
    faces = np.sort(np.concatenate([m.tets[:, [1, 2, 3]], m.tets[:, [0, 2, 3]], m.tets[:, [0, 1, 3]],
                                    m.tets[:, [0, 1, 2]]]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    assert set(counts.tolist()) <= {1, 2} and int((counts == 1).sum()) == 972


@pytest.mark.parametrize("shape", ["ellipsoid", "box"])
def test_solid_orientation_rest_shape(pkg, shape):
    m = pkg.scenes.solid(shape, 3000, seed=4)
    v = m.verts.astype(np.float64)
    a, b, c, d = (v[m.tets[:, i]] for i in range(4))
    det = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    assert (det > 0).all()                                          # positive orientation
    assert np.allclose(np.abs(det) / 6, m.tet_rest_volume, rtol=1e-5)
    Ds = np.stack([b - a, c - a, d - a], axis=2)
    F = Ds @ m.tet_inv_rest.astype(np.float64)                      # F = I at rest
    assert np.abs(F - np.eye(3)).max() < 1e-4
    assert np.allclose(np.linalg.norm(v[m.edges[:, 1]] - v[m.edges[:, 0]], axis=1), m.edge_rest, rtol=1e-6)
    assert abs(m.num_tets - 3000) / 3000 < 0.1
    assert float(m.verts[:, 1].min()) == 0.0
    # outward boundary normals: the surface encloses a positive volume
    p0, p1, p2 = (v[m.tris[:, i]] for i in range(3))
    enclosed = np.einsum("ij,ij->i", p0, np.cross(p1, p2)).sum() / 6.0
    assert np.isclose(enclosed, m.tet_rest_volume.sum(), rtol=1e-4)


def test_negative_orientation_option(pkg):
    m = pkg.scenes.cube_grid(5, orientation=-1)
    v = m.verts.astype(np.float64)
    a, b, c, d = (v[m.tets[:, i]] for i in range(4))
    assert (np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) < 0).all()


def test_mixed_sizes_distribution(pkg):
    sizes, grasp = pkg.scenes.mixed_sizes(1024, 1000, 40000, seed=0)
    t = np.array([s[1] for s in sizes])
    assert t.min() >= 1000 and t.max() <= 40000
    assert abs(np.log(t).mean() - 0.5 * (np.log(1000) + np.log(40000))) < 0.1   # log-uniform
    assert 0.4 < np.mean([s[0] == "box" for s in sizes]) < 0.6
    assert 0.4 < grasp.mean() < 0.6


def test_nearest_surface_vertices(pkg):
    m = pkg.scenes.solid("ellipsoid", 2000, seed=1)
    tip = pkg.scenes.top_point(m)
    g = pkg.scenes.nearest_surface_vertices(m, tip, 8)
    surf = set(m.surface_vertices().tolist())
    assert len(g) == 8 and all(int(i) in surf for i in g)
    d = np.linalg.norm(m.verts[g] - tip, axis=1)
    others = np.array(sorted(surf - set(g.tolist())))
    assert d.max() <= np.linalg.norm(m.verts[others] - tip, axis=1).min() + 1e-7


@pytest.mark.parametrize("which", ["tets", "edges"])
def test_coloring_is_conflict_free(pkg, which):
    m = pkg.scenes.solid("box", 5000, seed=2)
    topo = pkg.model.MeshTopology(m, constraint_order="color")
    elems = getattr(m, which)
    color = np.empty(len(elems), np.int32)
    nc = C.c_int32()
    pkg.lib.ffsrl_color_constraints(np.ascontiguousarray(elems, np.int32).ctypes.data, len(elems), elems.shape[1],
                                    m.num_vertices, color.ctypes.data, C.byref(nc))
    assert 1 < nc.value <= 64
    for c in range(nc.value):
        verts = elems[color == c].reshape(-1)
        assert len(np.unique(verts)) == len(verts), c              # no shared vertex inside a colour
    # the device order is the colour order, stable, and a permutation
    perm = topo.tet_perm if which == "tets" else topo.edge_perm
    assert sorted(perm.tolist()) == list(range(len(elems)))
    assert (np.diff(color[perm]) >= 0).all()
    dev = topo.tets if which == "tets" else topo.edges
    assert np.array_equal(dev, elems[perm])


def _lane_collisions(elems, lanes=64):
    """Vertex repeats inside groups of `lanes` consecutive constraints (the
    lanes of one wavefront's LDS atomic)."""
    n = 0
    for a in range(0, len(elems), lanes):
        v = elems[a:a + lanes].reshape(-1)
        n += len(v) - len(np.unique(v))
    return n


def _gather_lines(elems, lanes=64, per_line=8):
    """Distinct 128-byte lines of float4 positions touched per group of
    `lanes` consecutive constraints (the gathers of one wavefront)."""
    n = 0
    for a in range(0, len(elems), lanes):
        n += len(np.unique(elems[a:a + lanes].reshape(-1) // per_line))
    return n


@pytest.mark.parametrize("which", ["tets", "edges"])
def test_local_order(pkg, which):
    rng = np.random.default_rng(5)
    m = pkg.scenes.solid("ellipsoid", 6000, seed=4)
    elems = np.ascontiguousarray(getattr(m, which), np.int32)
    # scramble the element order so that locality must come from the ordering
    elems = elems[rng.permutation(len(elems))]
    verts = np.ascontiguousarray(m.verts, np.float32)

    def order(chunk):
        perm = np.empty(len(elems), np.int64)
        pkg.lib.ffsrl_order_constraints(elems.ctypes.data, len(elems), elems.shape[1], verts.ctypes.data,
                                        m.num_vertices, chunk, perm.ctypes.data)
        assert np.array_equal(np.sort(perm), np.arange(len(elems)))   # a permutation
        return perm

    local, morton = elems[order(128)], elems[order(1)]               # chunk 1: Morton order only
    # spatially local: far fewer distinct cache lines per wavefront than the scrambled order
    assert _gather_lines(local) < 0.5 * _gather_lines(elems)
    assert _gather_lines(morton) < 0.5 * _gather_lines(elems)
    # colour-sorted within runs of 128: fewer lanes of one atomic share a vertex
    assert _lane_collisions(local) < _lane_collisions(morton)
    # each run of 128 holds the same elements as the Morton order's run
    for a in range(0, len(elems), 128):
        assert sorted(map(tuple, local[a:a + 128])) == sorted(map(tuple, morton[a:a + 128]))
    perm = np.empty(len(elems), np.int64)
    # bad arguments are rejected
    with pytest.raises(pkg.FFSRLError):
        pkg.lib.ffsrl_order_constraints(elems.ctypes.data, len(elems), elems.shape[1], verts.ctypes.data,
                                        m.num_vertices, 0, perm.ctypes.data)
