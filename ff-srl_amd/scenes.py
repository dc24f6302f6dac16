"""Seeded synthetic tetrahedral scenes.

The reference ingests pre-tetrahedralised USD meshes whose ``extMesh:*``
attributes already hold edges, rest lengths, rest volumes and inverse rest
shape matrices (``FF_SRL/FF_SRL/modelDO.py:584-627``).  No USD reader is
available here, so scenes are generated: a voxel grid split into 5 tets per
hex cell with mirrored parity (so shared cell faces use the same diagonal and
the mesh is conforming), clipped to a box or an ellipsoid.  Everything the
reference reads from USD is derived here with the same meaning:

* ``edges`` - unique tet edges (``extMesh:edge``), rest length = |b - a|.
* ``tet_rest_volume`` - |det[b-a, c-a, d-a]| / 6 (``extMesh:tetrahedronRestVolume``).
* ``tet_inv_rest`` - inverse of the rest shape matrix whose COLUMNS are
  (b-a, c-a, d-a), stored row-major, so that ``F = Ds * Dm^-1`` is the identity
  at rest under the column-constructor reading of ``wp.mat33(v0, v1, v2)``
  used at ``FF_SRL/FF_SRL/integrator.py:627-629``.
* ``tris`` - boundary faces with outward normals (the render mesh; every
  render point is a simulation vertex, i.e. ``pointToVertex`` is the identity
  on the surface vertices, ``FF_SRL/FF_SRL/utils.py:32-42``).

Vertices are numbered in grid order and tets in cell order, which keeps the
vertices of neighbouring constraints close together in memory.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Corner c of a hex cell has offset (c & 1, (c >> 1) & 1, (c >> 2) & 1).
_CORNERS = np.array([[c & 1, (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=np.int64)
# Two mirrored 5-tet splits of a cube: the central tet uses the four corners of
# one parity, each corner tet a corner of the other parity plus its 3 neighbours.
_SPLIT_EVEN = np.array([[1, 2, 4, 7], [0, 1, 2, 4], [3, 1, 2, 7], [5, 1, 4, 7], [6, 2, 4, 7]], dtype=np.int64)
_SPLIT_ODD = np.array([[0, 3, 5, 6], [1, 0, 3, 5], [2, 0, 3, 6], [4, 0, 5, 6], [7, 3, 5, 6]], dtype=np.int64)


@dataclasses.dataclass
class TetMesh:
    """One simulation mesh in local vertex numbering (float32 / int32)."""
    name: str
    verts: np.ndarray            # (V, 3) float32 rest positions
    inv_mass: np.ndarray         # (V,) float32, 0 = pinned
    tets: np.ndarray             # (T, 4) int32
    tet_rest_volume: np.ndarray  # (T,) float32
    tet_inv_rest: np.ndarray     # (T, 3, 3) float32, row-major Dm^-1
    edges: np.ndarray            # (E, 2) int32, a < b
    edge_rest: np.ndarray        # (E,) float32
    tris: np.ndarray             # (F, 3) int32 boundary faces, outward
    colors: np.ndarray = None    # (V, 3) float32 vertex colours (render)
    tri_colors: np.ndarray = None  # (F, 3) float32 per-triangle colours
    uv: np.ndarray = None        # (V, 2) float32 texture coordinates of the vertices (texture mode)
    texture: np.ndarray = None   # (side, side, 3) float32 texels, or None (vertex colours are drawn)

    @property
    def num_vertices(self) -> int:
        return int(self.verts.shape[0])

    @property
    def num_tets(self) -> int:
        return int(self.tets.shape[0])

    @property
    def num_edges(self) -> int:
        return int(self.edges.shape[0])

    @property
    def num_tris(self) -> int:
        return int(self.tris.shape[0])

    def bbox_diagonal(self) -> float:
        ext = self.verts.max(axis=0).astype(np.float64) - self.verts.min(axis=0).astype(np.float64)
        return float(np.sqrt((ext * ext).sum()))

    def surface_vertices(self) -> np.ndarray:
        return np.unique(self.tris.reshape(-1)).astype(np.int32)


def _cells_to_mesh(name, cell_ijk, h, origin, jitter, rng, orientation=1):
    """Build a 5-tet mesh from a set of integer cell coordinates."""
    cell_ijk = np.asarray(cell_ijk, dtype=np.int64)
    parity = (cell_ijk.sum(axis=1) & 1).astype(bool)
    split = np.where(parity[:, None, None], _SPLIT_ODD[None], _SPLIT_EVEN[None])  # (C,5,4)
    corner_nodes = cell_ijk[:, None, :] + _CORNERS[None, :, :]                   # (C,8,3)
    tet_nodes = np.take_along_axis(corner_nodes[:, None, :, :].repeat(5, axis=1),
                                   split[..., None].repeat(3, axis=-1), axis=2)   # (C,5,4,3)
    tet_nodes = tet_nodes.reshape(-1, 4, 3)
    # Unique grid nodes in lexicographic (k, j, i) order -> local vertex ids.
    mx = corner_nodes.reshape(-1, 3).max(axis=0) + 1
    key = (tet_nodes[..., 2] * mx[1] + tet_nodes[..., 1]) * mx[0] + tet_nodes[..., 0]
    uniq, inv = np.unique(key.reshape(-1), return_inverse=True)
    tets = inv.reshape(-1, 4).astype(np.int64)
    node_ijk = np.stack([uniq % mx[0], (uniq // mx[0]) % mx[1], uniq // (mx[0] * mx[1])], axis=1)
    verts = origin[None, :] + node_ijk.astype(np.float64) * h
    if jitter > 0.0:
        verts = verts + rng.uniform(-jitter * h, jitter * h, size=verts.shape)
    # Orientation: det[b-a, c-a, d-a] has the requested sign.
    a, b, c, d = (verts[tets[:, i]] for i in range(4))
    det = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    flip = (det * orientation) < 0
    tets[flip] = tets[flip][:, [0, 2, 1, 3]]
    verts32 = verts.astype(np.float32)
    return name, verts32, tets.astype(np.int32)


def _finish(name, verts32, tets, inv_mass, rng=None) -> TetMesh:
    v64 = verts32.astype(np.float64)
    a, b, c, d = (v64[tets[:, i]] for i in range(4))
    ds = np.stack([b - a, c - a, d - a], axis=2)      # columns = edge vectors
    det = np.linalg.det(ds)
    rest_vol = (np.abs(det) / 6.0).astype(np.float32)
    inv_rest = np.linalg.inv(ds).astype(np.float32)   # row-major Dm^-1
    # Unique edges.
    pairs = tets[:, [0, 0, 0, 1, 1, 2, 1, 2, 3, 2, 3, 3]].reshape(-1, 2, 6).transpose(0, 2, 1).reshape(-1, 2)
    pairs = np.sort(pairs, axis=1)
    edges = np.unique(pairs, axis=0).astype(np.int32)
    edge_rest = np.linalg.norm(v64[edges[:, 1]] - v64[edges[:, 0]], axis=1).astype(np.float32)
    # Boundary faces: faces of tets that occur once; oriented outward.
    faces = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 3, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 1]]])
    opposite = np.concatenate([tets[:, 0], tets[:, 1], tets[:, 2], tets[:, 3]])
    skey = np.sort(faces, axis=1)
    _, first, counts = np.unique(skey, axis=0, return_index=True, return_counts=True)
    bnd = first[counts == 1]
    bnd.sort()
    tris = faces[bnd].copy()
    opp = opposite[bnd]
    p0, p1, p2 = v64[tris[:, 0]], v64[tris[:, 1]], v64[tris[:, 2]]
    n = np.cross(p1 - p0, p2 - p0)
    inward = np.einsum("ij,ij->i", n, v64[opp] - p0) > 0
    tris[inward] = tris[inward][:, [0, 2, 1]]
    rng = rng if rng is not None else np.random.default_rng(0)
    colors = rng.uniform(0.2, 1.0, size=(verts32.shape[0], 3)).astype(np.float32)
    tri_colors = rng.uniform(0.2, 1.0, size=(tris.shape[0], 3)).astype(np.float32)
    return TetMesh(name=name, verts=verts32, inv_mass=inv_mass.astype(np.float32),
                   tets=tets.astype(np.int32), tet_rest_volume=rest_vol, tet_inv_rest=inv_rest,
                   edges=edges, edge_rest=edge_rest, tris=tris.astype(np.int32),
                   colors=colors, tri_colors=tri_colors, uv=planar_uv(verts32), texture=seeded_texture(rng))


def planar_uv(verts: np.ndarray) -> np.ndarray:
    """Texture coordinates of a front (x, y) projection: s = (x - xmin) /
    (xmax - xmin), t = (y - ymin) / (ymax - ymin) of the rest positions."""
    v = np.asarray(verts, np.float64)
    lo, hi = v.min(axis=0), v.max(axis=0)
    ext = np.where(hi > lo, hi - lo, 1.0)
    return ((v[:, :2] - lo[:2]) / ext[:2]).astype(np.float32)


def seeded_texture(rng, side: int = 64, cells: int = 8) -> np.ndarray:
    """A side x side RGB texture: colours drawn on a coarse (cells + 1)^2 grid
    and interpolated bilinearly between its nodes, row-major (side, side, 3)."""
    grid = rng.uniform(0.2, 1.0, size=(cells + 1, cells + 1, 3))
    c = np.linspace(0.0, cells, side)
    i0 = np.minimum(np.floor(c).astype(np.int64), cells - 1)
    f = (c - i0)[:, None, None]
    rows = grid[i0] * (1.0 - f) + grid[i0 + 1] * f                        # (side, cells + 1, 3)
    g = (c - i0)[None, :, None]
    return (rows[:, i0] * (1.0 - g) + rows[:, i0 + 1] * g).astype(np.float32)


def cube_grid(n: int = 10, size: float = 1.0, orientation: int = 1, seed: int = 0) -> TetMesh:
    """n x n x n vertex grid on [0,size]^3, 5 tets per cell, bottom face (y=0) pinned."""
    cells = np.stack(np.meshgrid(np.arange(n - 1), np.arange(n - 1), np.arange(n - 1), indexing="ij"), -1).reshape(-1, 3)
    h = size / (n - 1)
    rng = np.random.default_rng(seed)
    name, verts, tets = _cells_to_mesh(f"cube{n}", cells, h, np.zeros(3), 0.0, rng, orientation)
    inv_mass = np.where(verts[:, 1] <= 0.0, 0.0, 1.0)
    return _finish(name, verts, tets, inv_mass, rng)


def _clip_cells(shape: str, semi_axes, h):
    a = np.asarray(semi_axes, dtype=np.float64)
    n = np.ceil(2.0 * a / h).astype(np.int64)
    cells = np.stack(np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij"), -1).reshape(-1, 3)
    origin = -0.5 * n.astype(np.float64) * h
    if shape == "box":
        return cells, origin
    centre = origin[None, :] + (cells.astype(np.float64) + 0.5) * h
    r = ((centre / a[None, :]) ** 2).sum(axis=1)
    return cells[r <= 1.0], origin


def _cells_for_target(shape, semi_axes, target_tets):
    a = np.asarray(semi_axes, dtype=np.float64)
    vol = (4.0 / 3.0) * math.pi * a.prod() if shape == "ellipsoid" else 8.0 * a.prod()
    h = (5.0 * vol / float(target_tets)) ** (1.0 / 3.0)
    # Refine h so the clipped cell count lands near the target, then keep the
    # closest count among nearby cell sizes (counts are quantised by the grid).
    for _ in range(6):
        cells, origin = _clip_cells(shape, a, h)
        got = 5 * len(cells)
        if got == 0:
            h *= 0.5
            continue
        ratio = got / float(target_tets)
        if abs(ratio - 1.0) < 0.02:
            break
        h *= ratio ** (1.0 / 3.0)
    best = (abs(5 * len(cells) - target_tets), cells, origin, h)
    if best[0] > 0.02 * target_tets:
        for f in np.linspace(0.9, 1.1, 21):
            c2, o2 = _clip_cells(shape, a, h * f)
            err = abs(5 * len(c2) - target_tets)
            if err < best[0]:
                best = (err, c2, o2, h * f)
    return best[1], best[2], best[3]


def solid(shape: str = "ellipsoid", target_tets: int = 20000, semi_axes=(1.0, 0.6, 0.4),
          pin_fraction: float = 0.10, jitter: float = 0.1, seed: int = 0,
          orientation: int = 1) -> TetMesh:
    """Seeded voxel-clipped ellipsoid or box with roughly ``target_tets`` tets.

    The mesh rests on the ground (min y = 0); the lowest ``pin_fraction`` of the
    vertices by height are pinned (inverse mass 0).
    """
    if shape not in ("ellipsoid", "box"):
        raise ValueError(f"unknown shape {shape!r}")
    rng = np.random.default_rng(seed)
    cells, origin, h = _cells_for_target(shape, semi_axes, target_tets)
    name, verts, tets = _cells_to_mesh(f"{shape}{5 * len(cells)}", cells, h, origin, jitter, rng, orientation)
    verts = (verts - np.array([0.0, verts[:, 1].min(), 0.0], dtype=np.float32)).astype(np.float32)
    order = np.argsort(verts[:, 1], kind="stable")
    npin = max(1, int(round(pin_fraction * verts.shape[0])))
    inv_mass = np.ones(verts.shape[0])
    inv_mass[order[:npin]] = 0.0
    return _finish(name, verts, tets, inv_mass, rng)


def nearest_surface_vertices(mesh: TetMesh, point, k: int = 8) -> np.ndarray:
    """The k surface vertices nearest ``point`` (grasp selection for the scenes)."""
    surf = mesh.surface_vertices()
    d = np.linalg.norm(mesh.verts[surf].astype(np.float64) - np.asarray(point, dtype=np.float64)[None], axis=1)
    return surf[np.argsort(d, kind="stable")[:k]].astype(np.int32)


def top_point(mesh: TetMesh) -> np.ndarray:
    """Highest surface vertex position: where the synthetic tool tip starts."""
    surf = mesh.surface_vertices()
    i = surf[np.argmax(mesh.verts[surf, 1])]
    return mesh.verts[i].astype(np.float32).copy()


def mixed_sizes(n_env: int, lo: int = 1000, hi: int = 40000, seed: int = 0):
    """Per-env (shape, target tets) with tets log-uniform in [lo, hi] and a 50/50 shape."""
    rng = np.random.default_rng(seed)
    tets = np.exp(rng.uniform(math.log(lo), math.log(hi), size=n_env))
    shapes = np.where(rng.uniform(size=n_env) < 0.5, "ellipsoid", "box")
    grasp = rng.uniform(size=n_env) < 0.5
    return [(str(s), int(round(t))) for s, t in zip(shapes, tets)], grasp
