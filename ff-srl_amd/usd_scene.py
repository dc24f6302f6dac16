"""Scenes of the reference's USD files: simulated tet meshes, lock boxes and
the laparoscope, read with the crate reader of ``usdc.py``.

This follows the reference's scene set-up:
* **Meshes.** A prim with ``extMesh:*`` attributes (``modelDO.py:545-640``)
  becomes a ``TetMesh``. Its vertices are mapped to world space by the
  prim's local-to-world transform (``model.py:47-48``). Tets, edges, rest
  lengths, rest volumes, inverse rest matrices, inverse masses and the
  surface triangles (``extMesh:triangle``) come from the file.
* **Lock boxes.** A ``Cube`` prim with ``simLockBox`` pins every vertex whose
  coordinates in the box's frame lie strictly inside (-1, 1) on all three
  axes (``modelDO.py:418-432``, ``:1664-1694``). The test is against 1, not
  the cube's half size of 0.5, as in the reference.
* **Laparoscope.** The prims flagged ``simLaparoscopeRotationCenter``,
  ``...Rod``, ``...RodCapsule``, ``...LeftClamp[RotationCenter]`` and
  ``...RightClamp[RotationCenter]`` give the seven transforms of
  ``laparoscope.py:302-386``. The capsule heights and radii come from the
  ``Capsule`` prims.

Transforms: a prim's local matrix is the product of its ``xformOpOrder`` ops
from first to last, in column-vector form. A ``rotateABC`` op turns about A
first, then B, then C, in degrees. Supported ops: translate, scale,
rotateX / Y / Z, the six rotateXYZ orders and transform (matrix4d).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

from .instrument import Laparoscope
from .scenes import TetMesh
from .usdc import CrateFile


def _rot(axis: int, deg: float) -> np.ndarray:
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    m = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def local_matrix(crate: CrateFile, prim: str) -> np.ndarray:
    """Local-to-parent matrix of a prim (column-vector form, float64)."""
    m = np.eye(4)
    for op in crate.get(prim, "xformOpOrder") or []:
        name = op.split(":")
        kind = name[1]
        v = crate.get(prim, op)
        if kind == "translate":
            t = np.eye(4)
            t[:3, 3] = np.asarray(v, np.float64)
        elif kind == "scale":
            t = np.diag([*np.asarray(v, np.float64), 1.0])
        elif kind in ("rotateX", "rotateY", "rotateZ"):
            t = _rot("XYZ".index(kind[-1]), float(v))
        elif kind.startswith("rotate") and len(kind) == 9:
            t = np.eye(4)
            for ax in kind[6:]:                      # first axis acts first
                k = "XYZ".index(ax)
                t = _rot(k, float(np.asarray(v)[k])) @ t
        elif kind == "transform":
            t = np.asarray(v, np.float64).T           # stored for row vectors
        else:
            raise ValueError(f"{prim}: unsupported transform op {op}")
        m = m @ t
    return m


def world_matrix(crate: CrateFile, prim: str) -> np.ndarray:
    """Local-to-world matrix: the ancestors' local matrices, root first."""
    parts = [p for p in prim.split("/") if p]
    m = np.eye(4)
    for k in range(1, len(parts) + 1):
        m = m @ local_matrix(crate, "/" + "/".join(parts[:k]))
    return m


def _points(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    p = np.asarray(p, np.float64)
    return p @ m[:3, :3].T + m[:3, 3]


def _flagged(crate: CrateFile, attr: str) -> List[str]:
    return [p for p in crate.prim_paths() if crate.get(p, attr) is True]


@dataclasses.dataclass
class UsdScene:
    meshes: List[TetMesh]
    params: List[Dict[str, float]]            # param:* of each mesh (modelDO.py:545-560)
    laparoscope: Optional[Laparoscope]
    meters_per_unit: float
    up_axis: str
    pinned: List[int]                          # vertices pinned by lock boxes, per mesh


def _mesh(crate: CrateFile, prim: str, boxes: List[np.ndarray]):
    g = lambda a: crate.get(prim, a)
    w = world_matrix(crate, prim)
    verts = _points(w, g("extMesh:vertex")).astype(np.float32)
    tets = np.asarray(g("extMesh:elem"), np.int32).reshape(-1, 4)
    edges = np.asarray(g("extMesh:edge"), np.int32).reshape(-1, 2)
    inv_mass = np.asarray(g("extMesh:inverseMass"), np.float32).copy()
    pinned = 0
    for box in boxes:                               # SimLockBoxDO.lockArray
        local = _points(np.linalg.inv(box), verts)
        inside = np.all((local < 1.0) & (local > -1.0), axis=1)
        pinned += int(np.count_nonzero(inside & (inv_mass != 0)))
        inv_mass[inside] = 0.0
    # Surface triangles: the render faces (faceVertexIndices over the render
    # points) mapped to simulation vertices through mapping:pointToVertex, which
    # is what the reference ray-casts after remapAToBKernel (utils.py:32-42).
    # They are extMesh:triangle in the same order with the opposite (outward)
    # winding; the collision only uses the triangle's plane, whose sign cancels.
    tris = np.asarray(g("extMesh:triangle"), np.int32).reshape(-1, 3)
    colors = np.full((verts.shape[0], 3), 0.6, np.float32)
    p2v, fvi, dc = g("mapping:pointToVertex"), g("faceVertexIndices"), g("primvars:displayColor")
    if p2v is not None and fvi is not None:
        p2v = np.asarray(p2v).reshape(-1, 2)
        point_vertex = np.full(int(p2v[:, 0].max()) + 1, -1, np.int64)
        point_vertex[p2v[:, 0]] = p2v[:, 1]
        faces = point_vertex[np.asarray(fvi).reshape(-1, 3)]
        if faces.shape == tris.shape and (faces >= 0).all():
            tris = faces.astype(np.int32)
        if dc is not None and len(dc) == len(point_vertex):   # vertex colours from the render points
            colors[point_vertex[point_vertex >= 0]] = np.asarray(dc, np.float32)[point_vertex >= 0]
    tri_colors = colors[tris].mean(axis=1).astype(np.float32)
    name = g("name") or prim.rsplit("/", 1)[-1]
    mesh = TetMesh(name=str(name), verts=verts, inv_mass=inv_mass, tets=tets,
                   tet_rest_volume=np.asarray(g("extMesh:tetrahedronRestVolume"), np.float32),
                   tet_inv_rest=np.asarray(g("extMesh:inverseRestPosition"), np.float32).reshape(-1, 3, 3),
                   edges=edges, edge_rest=np.asarray(g("extMesh:edgeRestLength"), np.float32),
                   tris=tris, colors=colors, tri_colors=tri_colors)
    params = {a.split(":", 1)[1]: float(crate.get(prim, a)) for a in crate.attributes(prim)
              if a.startswith("param:") and isinstance(crate.get(prim, a), (int, float))}
    return mesh, params, pinned


def _laparoscope(crate: CrateFile) -> Optional[Laparoscope]:
    roles = ["simLaparoscopeRodCapsule", "simLaparoscopeLeftClamp", "simLaparoscopeRightClamp",
             "simLaparoscopeRotationCenter", "simLaparoscopeLeftClampRotationCenter",
             "simLaparoscopeRightClampRotationCenter", "simLaparoscopeRod"]
    prims = []
    for r in roles:
        found = _flagged(crate, r)
        if not found:
            return None
        prims.append(found[0])
    xf = np.stack([world_matrix(crate, p) for p in prims])
    h = lambda p: float(crate.get(p, "height"))
    r = lambda p: float(crate.get(p, "radius"))
    return Laparoscope.from_transforms(xf, rod_height=h(prims[0]), rod_radius=r(prims[0]),
                                       left_height=h(prims[1]), left_radius=r(prims[1]),
                                       right_height=h(prims[2]), right_radius=r(prims[2]))


def load_usd_scene(path: str) -> UsdScene:
    """Simulated meshes, lock-box pins and the laparoscope of a USD crate
    scene (the reference's ``SimModelDO(stage, ...)`` inputs)."""
    crate = CrateFile(path)
    root = crate.fields_of("/")
    boxes = [world_matrix(crate, p) for p in _flagged(crate, "simLockBox")
             if crate.fields_of(p).get("typeName") == "Cube"]
    meshes, params, pinned = [], [], []
    for p in crate.prim_paths():
        if crate.get(p, "extMesh:elem") is not None:
            m, prm, n = _mesh(crate, p, boxes)
            meshes.append(m)
            params.append(prm)
            pinned.append(n)
    return UsdScene(meshes=meshes, params=params, laparoscope=_laparoscope(crate),
                    meters_per_unit=float(root.get("metersPerUnit") or 1.0), up_axis=str(root.get("upAxis") or "Y"),
                    pinned=pinned)


# ------------------------------------------------------------ converted scenes
_MESH_ARRAYS = ("verts", "inv_mass", "tets", "tet_rest_volume", "tet_inv_rest", "edges", "edge_rest", "tris",
                "colors", "tri_colors")


def save_scene(scene: UsdScene, path: str) -> None:
    """Write a loaded scene as a plain ``.npz`` (numbers only, no pickles),
    so a machine without the source ``.usd`` can load it."""
    out = {"n_meshes": np.array(len(scene.meshes)), "meters_per_unit": np.array(scene.meters_per_unit),
           "up_axis": np.array(scene.up_axis), "pinned": np.asarray(scene.pinned, np.int64)}
    for i, m in enumerate(scene.meshes):
        for a in _MESH_ARRAYS:
            out[f"mesh{i}_{a}"] = getattr(m, a)
        out[f"mesh{i}_name"] = np.array(m.name)
        keys = sorted(scene.params[i])
        out[f"mesh{i}_param_names"] = np.array(keys)
        out[f"mesh{i}_param_values"] = np.array([scene.params[i][k] for k in keys], np.float64)
    lap = scene.laparoscope
    if lap is not None:
        out["lap_xf"] = lap.transforms()
        out["lap_sizes"] = np.array([lap.rod_height, lap.rod_radius, *lap.left_jaw, *lap.right_jaw], np.float64)
    np.savez_compressed(path, **out)


def load_scene(path: str) -> UsdScene:
    """Read a scene written by ``save_scene`` (``numpy.load`` without pickles)."""
    z = np.load(path, allow_pickle=False)
    meshes, params = [], []
    for i in range(int(z["n_meshes"])):
        arr = {a: z[f"mesh{i}_{a}"] for a in _MESH_ARRAYS}
        meshes.append(TetMesh(name=str(z[f"mesh{i}_name"]), **arr))
        params.append(dict(zip((str(k) for k in z[f"mesh{i}_param_names"]),
                               (float(v) for v in z[f"mesh{i}_param_values"]))))
    lap = None
    if "lap_xf" in z:
        s = z["lap_sizes"]
        lap = Laparoscope.from_transforms(z["lap_xf"], s[0], s[1], s[2], s[3], s[4], s[5])
    return UsdScene(meshes=meshes, params=params, laparoscope=lap, meters_per_unit=float(z["meters_per_unit"]),
                    up_axis=str(z["up_axis"]), pinned=[int(v) for v in z["pinned"]])


def merged_mesh(scene: UsdScene) -> TetMesh:
    """The scene's simulated meshes as one mesh (an environment of the batch
    holds one mesh; constraints never join two meshes, so merging changes
    nothing)."""
    if len(scene.meshes) == 1:
        return scene.meshes[0]
    base = np.cumsum([0] + [m.num_vertices for m in scene.meshes])
    cat = lambda a: np.concatenate([getattr(m, a) for m in scene.meshes])
    shift = lambda a: np.concatenate([getattr(m, a) + base[i] for i, m in enumerate(scene.meshes)]).astype(np.int32)
    return TetMesh(name="+".join(m.name for m in scene.meshes), verts=cat("verts"), inv_mass=cat("inv_mass"),
                   tets=shift("tets"), tet_rest_volume=cat("tet_rest_volume"), tet_inv_rest=cat("tet_inv_rest"),
                   edges=shift("edges"), edge_rest=cat("edge_rest"), tris=shift("tris"), colors=cat("colors"),
                   tri_colors=cat("tri_colors"))


def scene(path: str) -> UsdScene:
    """A ``.usd`` / ``.usdc`` crate file or a converted ``.npz``."""
    return load_scene(path) if path.endswith(".npz") else load_usd_scene(path)
