"""SimModelDO: data-oriented multi-environment state on the GPU.

Mirrors ``FF_SRL/FF_SRL/modelDO.py:772-1645`` (``SimModelDO``): the state of
all environments lives in flat arrays, every kernel processes all of them, and
observations are torch tensors.  Differences by design:

* Environments may use different meshes.  A CSR environment table (one
  ``ffsrl_env_desc`` per env) gives each env its state range and the ranges of
  its mesh's shared topology (edges, tets, render triangles, BVH nodes), so
  identical envs share one copy of the topology instead of the reference's
  per-env replicas (``modelDO.py:916-999``).
* Positions are float4 (x, y, z, inverse mass) so a constraint gathers a
  vertex and its mass in one 16-byte load.
* The BVH topology of each mesh is built here (host binned SAH in the C++
  library) because the batched frame kernel refits it in the same launch.
"""
from __future__ import annotations

import concurrent.futures
import ctypes as C
import math
import os
from typing import Optional, Sequence

import numpy as np
import torch

from ._lib import (Batch, BvhJob, EnvDesc, ENV_DESC_FIELDS, FFSRLError, Instrument, OFF, PART_OWNER, PART_SHARED,
                   PBD, XPBD, Params, Parts, STATUS_INCOMPLETE, STATUS_SATURATED, lib)
from .scenes import TetMesh

MAX_STACK = 32  # traversal stack entries per ray (render.py:424, 924)


def _pack16(ids: np.ndarray) -> np.ndarray:
    """Rows of 4 (tets) or 2 (edges) mesh-local ids packed as 16-bit fields,
    first id in the low bits: int64 [n] for 4 ids, int32 [n] for 2.  Ids of
    meshes with more than 65,536 vertices are stored modulo 2^16 (such meshes
    exceed the batched frame's 10,208-vertex limit and use the plain form)."""
    ids = np.asarray(ids).astype(np.uint64) & np.uint64(0xffff)
    if ids.ndim != 2 or ids.shape[1] not in (2, 4):
        raise ValueError("expected (n, 2) or (n, 4) ids")
    out = np.zeros(ids.shape[0], np.uint64)
    for k in range(ids.shape[1]):
        out |= ids[:, k] << np.uint64(16 * k)
    return out.view(np.int64) if ids.shape[1] == 4 else out.astype(np.uint32).view(np.int32)


def _parallel(fn, items):
    """Map over host threads: the library's host routines (BVH build,
    constraint order) release the interpreter lock while they run."""
    items = list(items)
    if len(items) < 8:
        return [fn(x) for x in items]
    lib.load()
    with concurrent.futures.ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(fn, items))
LOCAL_CHUNK = 128  # constraints per colour-sorted run of the "local" order
PART_TETS = 5120   # constraints per part of the partitioned frame (tets, or edges without tet constraints)
PART_MAX = 32      # most parts per mesh (one XCD's CUs run an env's parts at once)


def partition_mesh(tets: np.ndarray, edges: np.ndarray, nv: int, parts: int, use_tets: bool = True,
                   use_edges: bool = True, tris: Optional[np.ndarray] = None) -> dict:
    """Cuts a mesh's constraints (in their current order) into ``parts``
    ranges of tets and of edges for the partitioned frame
    (``ffsrl_partition_mesh``).  Returns the part rows (tet begin / end, edge
    begin / end, first vertex entry, vertex count, first tris16 entry,
    triangle count; entries relative to this mesh), each part's vertices (mesh-local id |
    PART_SHARED when several parts hold it | PART_OWNER on the lowest part
    holding it), the part-local ids of every tet and edge (packed as the
    frame reads them: 4 x 16 bits, 2 x 16 bits), and the largest part.
    With ``tris`` (the surface triangles, int4 in BVH leaf order: the
    collision pass), each triangle goes to the lowest part holding its first
    vertex, its vertices join that part, and ``tris16`` holds every part's
    triangles' local ids (3 x 16 bits), by part.  Only the constraint
    families in use (and the triangles) decide a part's vertices; a vertex
    that none of them touches goes to part i mod parts (i its rank among
    such vertices)."""
    P = int(parts)
    tets = np.ascontiguousarray(np.asarray(tets).reshape(-1, 4), dtype=np.int32)
    edges = np.ascontiguousarray(np.asarray(edges).reshape(-1, 2), dtype=np.int32)
    use_tris = tris is not None
    tris = np.ascontiguousarray(np.asarray(tris if use_tris else np.zeros((0, 4))).reshape(-1, 4), dtype=np.int32)
    nt, ne, nf = tets.shape[0], edges.shape[0], tris.shape[0]
    rows = np.zeros((P, 8), np.int32)
    cap = 4 * nt + 2 * ne + 3 * nf + nv
    verts = np.zeros(max(cap, 1), np.int32)
    t16 = np.zeros(max(nt, 1), np.uint64)
    e16 = np.zeros(max(ne, 1), np.uint32)
    f16 = np.zeros(max(nf, 1), np.uint64)
    n_verts, most = C.c_int64(0), C.c_int32(0)
    lib.ffsrl_partition_mesh(tets.ctypes.data, nt, edges.ctypes.data, ne, tris.ctypes.data, nf, int(nv), P,
                             int(bool(use_tets)), int(bool(use_edges)), int(use_tris), rows.ctypes.data,
                             verts.ctypes.data, cap, C.byref(n_verts), t16.ctypes.data, e16.ctypes.data,
                             f16.ctypes.data, C.byref(most))
    return {"rows": rows.astype(np.int64), "verts": verts[:n_verts.value], "tets16": t16[:nt].view(np.int64),
            "edges16": e16[:ne].view(np.int32), "tris16": f16[:nf].view(np.int64), "max_local": int(most.value),
            "parts": P}

_MODES = {None: OFF, "off": OFF, "pbd": PBD, "xpbd": XPBD}


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def surface_inputs(mesh: TetMesh):
    """The BVH build's inputs: the surface vertices (the render points) and the
    render faces in their numbering."""
    surf = mesh.surface_vertices()
    local = np.full(mesh.num_vertices, -1, dtype=np.int32)
    local[surf] = np.arange(surf.size, dtype=np.int32)
    faces_s = np.ascontiguousarray(local[mesh.tris]).astype(np.int32)
    verts_s = np.ascontiguousarray(mesh.verts[surf]).astype(np.float32)
    return verts_s, faces_s


def build_bvh_host(mesh: TetMesh, bins: int = 8) -> dict:
    """SAH topology of one mesh on the host (``ffsrl_bvh_build``)."""
    verts_s, faces_s = surface_inputs(mesh)
    nf = mesh.num_tris
    out = {k: np.empty(2 * nf, np.int32) for k in ("left", "first", "count", "level")}
    out["tri_map"] = np.empty(nf, np.int32)
    maxl, nn = C.c_int32(0), C.c_int32(0)
    lib.ffsrl_bvh_build(verts_s.ctypes.data, verts_s.shape[0], faces_s.ctypes.data, nf, bins,
                        out["left"].ctypes.data, out["first"].ctypes.data, out["count"].ctypes.data,
                        out["level"].ctypes.data, out["tri_map"].ctypes.data, C.byref(nn), C.byref(maxl))
    out["num_nodes"], out["max_level"] = int(nn.value), int(maxl.value)
    return out


def build_bvhs_device(meshes: Sequence[TetMesh], device="cuda:0", bins: int = 8) -> list:
    """SAH topologies of many meshes in one launch (``ffsrl_bvh_build_device``),
    node for node the host build's.  Returns one dict per mesh with the arrays
    MeshTopology takes."""
    dev = torch.device(device)
    inputs = [surface_inputs(m) for m in meshes]
    sizes = [(v.shape[0], f.shape[0]) for v, f in inputs]
    align = lambda x: (x + 255) // 256 * 256
    scratch_sz = [align(int(lib.ffsrl_bvh_scratch_bytes(nf))) for _, nf in sizes]
    scratch = torch.empty(max(sum(scratch_sz), 1), dtype=torch.uint8, device=dev)
    verts = [torch.from_numpy(v.reshape(-1)).to(dev) for v, _ in inputs]
    faces = [torch.from_numpy(f.reshape(-1)).to(dev) for _, f in inputs]
    outs = [torch.empty(8 * nf + nf + 2, dtype=torch.int32, device=dev) for _, nf in sizes]
    jobs = (BvhJob * len(meshes))()
    off = 0
    for i, ((nv, nf), o) in enumerate(zip(sizes, outs)):
        j = jobs[i]
        j.verts, j.faces, j.nv, j.nf = verts[i].data_ptr(), faces[i].data_ptr(), nv, nf
        base = o.data_ptr()
        j.left, j.first, j.count, j.level = base, base + 8 * nf, base + 16 * nf, base + 24 * nf
        j.tri_map, j.info = base + 32 * nf, base + 36 * nf
        j.scratch = scratch.data_ptr() + off
        off += scratch_sz[i]
    jobs_dev = torch.frombuffer(bytearray(jobs), dtype=torch.uint8).to(dev)
    lib.ffsrl_bvh_build_device(jobs_dev.data_ptr(), len(meshes), int(bins), torch.cuda.current_stream().cuda_stream)
    res = []
    for (nv, nf), o in zip(sizes, outs):
        a = o.cpu().numpy()
        n, max_level = int(a[9 * nf]), int(a[9 * nf + 1])
        if n < 0:
            raise ValueError("device BVH build: tree deeper than 64 levels")
        res.append({"left": a[:2 * nf], "first": a[2 * nf:4 * nf], "count": a[4 * nf:6 * nf],
                    "level": a[6 * nf:8 * nf], "tri_map": a[8 * nf:9 * nf].copy(), "num_nodes": n,
                    "max_level": max_level})
    return res


def validate_bvh(bvh: dict, nf: int) -> None:
    """Raises ValueError unless ``bvh`` is a tree the kernels can traverse
    and refit safely: the root is node 0 at level 0; an inner node (count
    0) has its children at left and left + 1, both in range and one level
    deeper; a leaf (count > 0, left -1) holds triangles [first, first +
    count) within [0, nf); every node is reached from the root exactly once;
    the leaves' ranges tile [0, nf) exactly once; tri_map permutes the nf
    triangles.  The render and refit kernels index nodes and triangles with
    these values unchecked."""
    n = int(bvh["num_nodes"])
    left, first, count, level = (np.asarray(bvh[k], np.int64) for k in ("left", "first", "count", "level"))
    if n < 1 or min(left.size, first.size, count.size, level.size) < n:
        raise ValueError(f"BVH: {n} nodes declared, arrays hold {min(left.size, first.size, count.size, level.size)}")
    if n > max(1, 2 * nf - 1):
        raise ValueError(f"BVH: {n} nodes for {nf} triangles (at most {2 * nf - 1})")
    tri_map = np.asarray(bvh["tri_map"])
    if tri_map.shape != (nf,) or not np.array_equal(np.sort(tri_map), np.arange(nf)):
        raise ValueError(f"BVH: the triangle map does not permute the {nf} triangles")
    if level[0] != 0:
        raise ValueError("BVH: the root is not at level 0")
    seen = np.zeros(n, bool)
    cover = np.zeros(nf, np.int64)
    stack = [0]
    while stack:
        k = stack.pop()
        if seen[k]:
            raise ValueError(f"BVH: node {k} is reached twice")
        seen[k] = True
        if count[k] > 0:
            if left[k] != -1:
                raise ValueError(f"BVH: leaf {k} has a left child")
            if first[k] < 0 or first[k] + count[k] > nf:
                raise ValueError(f"BVH: leaf {k} holds triangles [{first[k]}, {first[k] + count[k]}) of {nf}")
            cover[first[k]:first[k] + count[k]] += 1
        elif count[k] == 0:
            a = left[k]
            if not (0 < a and a + 1 < n):
                raise ValueError(f"BVH: inner node {k} has children {a}, {a + 1} outside (0, {n})")
            if level[a] != level[k] + 1 or level[a + 1] != level[k] + 1:
                raise ValueError(f"BVH: the children of node {k} are not one level deeper")
            stack += [int(a), int(a + 1)]
        else:
            raise ValueError(f"BVH: node {k} has a negative triangle count")
    if not seen.all():
        raise ValueError(f"BVH: nodes {np.nonzero(~seen)[0][:5].tolist()} are not reached from the root")
    if (cover != 1).any():
        raise ValueError("BVH: the leaves do not hold every triangle exactly once")
    if int(bvh["max_level"]) != int(level[:n].max()):
        raise ValueError("BVH: max_level differs from the deepest node level")


class MeshTopology:
    """Host-side topology of one mesh plus its BVH (built once)."""

    def __init__(self, mesh: TetMesh, bins: int = 8, constraint_order: str = "local", chunk: int = LOCAL_CHUNK,
                 bvh: Optional[dict] = None):
        """`bvh`: the SAH topology already built (``build_bvhs_device``);
        otherwise it is built here on the host (``ffsrl_bvh_build``)."""
        self.mesh = mesh
        nf = mesh.num_tris
        if bvh is None:
            bvh = build_bvh_host(mesh, bins)
        left, first, count, level, tri_map = (bvh[k] for k in ("left", "first", "count", "level", "tri_map"))
        n, max_level = int(bvh["num_nodes"]), int(bvh["max_level"])
        self.num_nodes = n
        self.max_level = max_level
        if self.max_level + 1 > MAX_STACK:
            raise ValueError(f"BVH of {mesh.name} has {self.max_level + 1} levels; the ray stack holds {MAX_STACK}")
        self.left = left[:n].copy()
        self.first = first[:n].copy()
        self.count = count[:n].copy()
        self.level = level[:n].copy()
        self.tri_map = tri_map
        # Unsplit root: the reference marks it with leftNode 0 (bvh.py:480); -1 is used here.
        # Leaf-order triangles: (a, b, c, original face id).
        self.tris_leaf = np.concatenate([mesh.tris[tri_map], tri_map[:, None]], axis=1).astype(np.int32)
        # Refit groups: every leaf first (leaves depend only on positions), then
        # the inner nodes level by level, deepest first.  level_info holds the
        # node's (left child or first triangle, triangle count) per entry.
        leaf = self.count > 0
        order = [np.nonzero(leaf)[0]]
        self.level_off = [0, order[0].size]
        for L in range(self.max_level, -1, -1):
            ids = np.nonzero((self.level == L) & ~leaf)[0]
            if ids.size:
                order.append(ids)
                self.level_off.append(self.level_off[-1] + ids.size)
        self.level_nodes = np.concatenate(order).astype(np.int32)
        self.level_off = np.asarray(self.level_off, dtype=np.int32)
        self.num_groups = self.level_off.size - 1
        nodes = self.level_nodes
        self.level_info = np.stack([np.where(leaf[nodes], self.first[nodes], self.left[nodes]),
                                    self.count[nodes]], axis=1).astype(np.int32)
        self.tri_colors_leaf = mesh.tri_colors[tri_map] if mesh.tri_colors is not None else None
        # texture coordinates of each leaf-order triangle's corners (s, t) x 3
        self.tri_uv_leaf = (np.ascontiguousarray(mesh.uv[mesh.tris[tri_map]].reshape(nf, 6)).astype(np.float32)
                            if mesh.uv is not None and mesh.texture is not None else None)
        # vertex -> faces adjacency for the normals: faces of each vertex in
        # increasing original face id, stored as leaf positions
        leaf_of = np.empty(nf, np.int64)
        leaf_of[tri_map] = np.arange(nf)
        corner_v = mesh.tris.reshape(-1).astype(np.int64)
        order_vf = np.argsort(corner_v, kind="stable")            # face ids stay increasing per vertex
        self.vf_idx = leaf_of[order_vf // 3].astype(np.int32)
        self.vf_off = np.zeros(mesh.num_vertices + 1, np.int32)
        np.cumsum(np.bincount(corner_v, minlength=mesh.num_vertices), out=self.vf_off[1:])
        self.set_order(constraint_order, chunk)

    def set_order(self, constraint_order: str = "local", chunk: int = LOCAL_CHUNK) -> None:
        """Constraint order on the device (see ``_order``); the frame's integer
        Jacobi sums make results independent of it."""
        mesh = self.mesh
        self.constraint_order, self.chunk = constraint_order, int(chunk)
        self.tet_perm = self._order(mesh.tets, mesh, constraint_order, chunk)
        self.edge_perm = self._order(mesh.edges, mesh, constraint_order, chunk)
        self.tets = np.ascontiguousarray(mesh.tets[self.tet_perm]).astype(np.int32)
        self.tet_rest_volume = np.ascontiguousarray(mesh.tet_rest_volume[self.tet_perm])
        self.tet_inv_rest = np.ascontiguousarray(mesh.tet_inv_rest[self.tet_perm].reshape(-1, 9))
        self.edges = np.ascontiguousarray(mesh.edges[self.edge_perm]).astype(np.int32)
        self.edge_rest = np.ascontiguousarray(mesh.edge_rest[self.edge_perm])

    @staticmethod
    def _order(elems: np.ndarray, mesh: TetMesh, order: str, chunk: int = LOCAL_CHUNK) -> np.ndarray:
        """Visit order of the constraints in the frame kernel:
        "local" - Morton order of the centroids, greedy-coloured within runs of
                  `chunk` (position gathers of a wavefront hit few cache lines
                  and its LDS atomics rarely collide);
        "color" - stable sort by a greedy colour of the whole mesh;
        "mesh"  - the mesh's own order."""
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        n = elems.shape[0]
        if n == 0 or order == "mesh":
            return np.arange(n, dtype=np.int64)
        if order == "color":
            color = np.empty(n, np.int32)
            nc = C.c_int32(0)
            lib.ffsrl_color_constraints(elems.ctypes.data, n, elems.shape[1], mesh.num_vertices, color.ctypes.data,
                                        C.byref(nc))
            return np.argsort(color, kind="stable")
        if order != "local":
            raise ValueError(f"unknown constraint order {order!r}")
        verts = np.ascontiguousarray(mesh.verts, dtype=np.float32)
        perm = np.empty(n, np.int64)
        lib.ffsrl_order_constraints(elems.ctypes.data, n, elems.shape[1], verts.ctypes.data, mesh.num_vertices,
                                    int(chunk), perm.ctypes.data)
        return perm


class SimModelDO:
    """Batched environments (``modelDO.py:772``) on one GPU.

    ``meshes[env_mesh[i]]`` is the mesh of env ``i``.  Physics parameters keep
    the reference's constructor names (``modelDO.py:774-802``); the constraint
    forms are chosen with ``distance`` / ``volume`` ("pbd", "xpbd" or None) and
    ``neoHookean``.  XPBD compliances are given as alpha and converted to
    alpha~ = alpha / dt^2 (``modelDO.py:828-830``).
    """

    def __init__(self, meshes: Sequence[TetMesh], env_mesh: Optional[Sequence[int]] = None, device="cuda:0",
                 simSubsteps=8, simFrameRate=60, simConstraintsSteps=1,
                 environmentGravity=(0.0, -9.8, 0.0), environmentVelocityDampening=0.1,
                 environmentGroundLevel=0.0, globalKsDistance=1.0, globalKsVolume=1.0, globalKsDrag=1.0,
                 distance="pbd", volume="pbd", neoHookean=False, distanceCompliance=0.0,
                 volumeComplianceXPBD=0.0, globalVolumeCompliance=1.0, globalDeviatoricCompliance=1.0,
                 globalLaparoscopeDragLookupRadius=0.15, lambdaLegacy=False, bins=8,
                 initialVelocity: Optional[np.ndarray] = None, tips: Optional[np.ndarray] = None,
                 constraint_order: str = "local", globalKsContact=1.0, collision: bool = False,
                 capsuleBases: Optional[np.ndarray] = None, capsuleRadii: Optional[np.ndarray] = None,
                 laparoscopes=None, workspaceLow=(-1e5, -1e5, -1e5), workspaceHigh=(1e5, 1e5, 1e5),
                 bvh_on_device: bool = False):
        """``meshes`` and ``env_mesh``: the simulated meshes and each env's
        mesh.  The reference's form ``SimModelDO(stage, numEnvs, device, ...)``
        (``modelDO.py:774``) is accepted too: a USD crate scene path (or its
        converted ``.npz``), or an opened stage (its root layer's file), and
        the number of envs; every env then holds the scene's meshes and
        laparoscope (``self.scene`` keeps the parsed scene)."""
        self.scene = None
        if isinstance(meshes, (str, os.PathLike)) or hasattr(meshes, "GetRootLayer"):
            from .usd_scene import merged_mesh, scene
            path = meshes.GetRootLayer().realPath if hasattr(meshes, "GetRootLayer") else os.fspath(meshes)
            self.scene = scene(path)
            n = int(env_mesh)
            if laparoscopes is None and self.scene.laparoscope is not None:
                laparoscopes = [self.scene.laparoscope] * n
            meshes, env_mesh = [merged_mesh(self.scene)], [0] * n
        self.device = torch.device(device)
        self.meshes = list(meshes)
        env_mesh = list(range(len(self.meshes))) if env_mesh is None else list(env_mesh)
        self.env_mesh = np.asarray(env_mesh, dtype=np.int64)
        self.numEnvs = len(env_mesh)
        self.simSubsteps = int(simSubsteps)
        self.simFrameRate = simFrameRate
        self.simConstraints = int(simConstraintsSteps)
        self.simDt = (1.0 / self.simFrameRate) / self.simSubsteps
        self.gravity = tuple(float(g) for g in environmentGravity)
        self.velocityDampening = environmentVelocityDampening
        self.groundLevel = environmentGroundLevel
        self.globalKsDistance = globalKsDistance
        self.globalKsVolume = globalKsVolume
        self.globalKsDrag = globalKsDrag
        self.globalLaparoscopeDragLookupRadius = globalLaparoscopeDragLookupRadius
        self.dist_mode = _MODES[distance]
        self.vol_mode = _MODES[volume]
        self.nh_on = bool(neoHookean)
        self.distanceCompliance = distanceCompliance
        self.volumeComplianceXPBD = volumeComplianceXPBD
        self.globalVolumeComplianceNeoHookean = globalVolumeCompliance
        self.gloabalDeviatoricComplianceNeoHookean = globalDeviatoricCompliance
        self.lambda_legacy = bool(lambdaLegacy)
        self.globalKsContact = globalKsContact
        self.collision = bool(collision)
        self.workspaceLow = tuple(float(v) for v in workspaceLow)
        self.workspaceHigh = tuple(float(v) for v in workspaceHigh)
        self.params = self._make_params()
        if laparoscopes is not None:
            # instrument points, capsules and heights from the laparoscope geometry
            laps = list(laparoscopes) if isinstance(laparoscopes, (list, tuple)) else [laparoscopes] * len(env_mesh)
            if len(laps) != len(env_mesh):
                raise ValueError(f"{len(laps)} laparoscopes for {len(env_mesh)} environments")
            pts = [lap.initial_points() for lap in laps]
            tips = np.stack([p[0] for p in pts]).astype(np.float32)
            capsuleBases = np.stack([p[1] for p in pts]).astype(np.float32)
            capsuleRadii = np.stack([p[2] for p in pts]).astype(np.float32)

        self.constraint_order = constraint_order
        if bvh_on_device:   # every mesh's SAH topology in one device launch
            bvhs = build_bvhs_device(self.meshes, self.device, bins)
            self.topo = _parallel(lambda mb: MeshTopology(mb[0], bins, constraint_order, bvh=mb[1]),
                                  list(zip(self.meshes, bvhs)))
        else:
            self.topo = _parallel(lambda m: MeshTopology(m, bins, constraint_order), self.meshes)
        self._build_topology()
        self._build_envs()
        self._build_state(initialVelocity, tips, capsuleBases, capsuleRadii)
        self._batch = self._make_batch()
        self._instrument = None
        if laparoscopes is not None:
            self._build_instrument(laps)
        self._order = None

    @classmethod
    def from_usd(cls, path: str, numEnvs: int, device="cuda:0", **kw) -> "SimModelDO":
        """``SimModelDO(stage, numEnvs, device, ...)`` of the reference
        (``modelDO.py:774-1266``) for a USD crate scene or its converted
        ``.npz``: every env holds the scene's simulated meshes (merged into one
        mesh, lock-box pins applied) and its laparoscope."""
        return cls(path, int(numEnvs), device=device, **kw)

    # ------------------------------------------------------------------ setup
    def _make_params(self) -> Params:
        p = Params()
        p.substeps = self.simSubsteps
        p.iterations = self.simConstraints
        p.dt = self.simDt
        p.gravity[:] = self.gravity
        p.damping = self.velocityDampening
        p.ground = self.groundLevel
        p.dist_mode = self.dist_mode
        p.dist_k = self.globalKsDistance if self.dist_mode == PBD else self.distanceCompliance / (self.simDt * self.simDt)
        p.vol_mode = self.vol_mode
        p.vol_k = self.globalKsVolume if self.vol_mode == PBD else self.volumeComplianceXPBD / (self.simDt * self.simDt)
        p.nh_on = int(self.nh_on)
        p.nh_vol_c = self.globalVolumeComplianceNeoHookean
        p.nh_dev_c = self.gloabalDeviatoricComplianceNeoHookean
        p.drag_ks = self.globalKsDrag
        p.lambda_legacy = int(self.lambda_legacy)
        p.collide = int(self.collision)
        p.contact_k = self.globalKsContact
        return p

    @property
    def carries_lambda(self) -> bool:
        xpbd = self.dist_mode == XPBD or self.vol_mode == XPBD
        return xpbd and (self.simConstraints > 1 or self.lambda_legacy)

    def _build_topology(self):
        dev = self.device
        off = {"e": [0], "t": [0], "f": [0], "n": [0], "l": [0], "vt": [0]}
        for t in self.topo:
            m = t.mesh
            off["e"].append(off["e"][-1] + m.num_edges)
            off["t"].append(off["t"][-1] + m.num_tets)
            off["f"].append(off["f"][-1] + m.num_tris)
            off["n"].append(off["n"][-1] + t.num_nodes)
            off["l"].append(off["l"][-1] + t.level_off.size)
            off["vt"].append(off["vt"][-1] + t.vf_off.size)
        self.mesh_off = {k: np.asarray(v, dtype=np.int64) for k, v in off.items()}
        cat = np.concatenate
        self.t_edges = torch.from_numpy(cat([t.edges for t in self.topo])).to(dev)
        self.t_edge_rest = torch.from_numpy(cat([t.edge_rest for t in self.topo])).to(dev)
        self.t_tets = torch.from_numpy(cat([t.tets for t in self.topo])).to(dev)
        self.t_tets16 = torch.from_numpy(_pack16(cat([t.tets for t in self.topo]))).to(dev)
        self.t_edges16 = None
        self.t_tet_rest_volume = torch.from_numpy(cat([t.tet_rest_volume for t in self.topo])).to(dev)
        self.t_edges16 = torch.from_numpy(_pack16(cat([t.edges for t in self.topo]))).to(dev)
        inv_rest = cat([t.tet_inv_rest for t in self.topo]).astype(np.float32)
        self.t_tet_inv_rest = torch.from_numpy(np.ascontiguousarray(inv_rest.T)).to(dev)
        self.t_tris = torch.from_numpy(cat([t.tris_leaf for t in self.topo])).to(dev)
        self.t_node_left = torch.from_numpy(cat([t.left for t in self.topo])).to(dev)
        self.t_node_first = torch.from_numpy(cat([t.first for t in self.topo])).to(dev)
        self.t_node_count = torch.from_numpy(cat([t.count for t in self.topo])).to(dev)
        self.t_level_nodes = torch.from_numpy(cat([t.level_nodes for t in self.topo])).to(dev)
        self.t_level_info = torch.from_numpy(cat([t.level_info for t in self.topo])).to(dev)
        self.t_vf_off = torch.from_numpy(cat([t.vf_off for t in self.topo])).to(dev)
        self.t_vf_idx = torch.from_numpy(cat([t.vf_idx for t in self.topo])).to(dev)
        self.level_off_host = np.ascontiguousarray(cat([t.level_off for t in self.topo]).astype(np.int32))
        self.t_level_off = torch.from_numpy(self.level_off_host).to(dev)
        tc = [t.tri_colors_leaf if t.tri_colors_leaf is not None else np.ones((t.mesh.num_tris, 3), np.float32)
              for t in self.topo]
        self.t_tri_colors = torch.from_numpy(cat(tc).astype(np.float32)).to(dev)
        self._build_textures()
        self.n_tet_total = int(self.t_tets.shape[0])

    def _build_textures(self):
        """Texture table of the render meshes (colour mode 1): the texels of
        every mesh that has a texture (float4, each texture's consecutive),
        (first texel, side) per texture, and per leaf-order triangle its
        corners' texture coordinates and texture id (-1: vertex colours)."""
        dev = self.device
        uv, tex_id, texels, info, first = [], [], [], [], 0
        for t in self.topo:
            m, nf = t.mesh, t.mesh.num_tris
            if t.tri_uv_leaf is None:
                uv.append(np.zeros((nf, 6), np.float32))
                tex_id.append(np.full(nf, -1, np.int32))
                continue
            side = int(m.texture.shape[0])
            tx = np.zeros((side * side, 4), np.float32)
            tx[:, :3] = m.texture.reshape(-1, 3)
            uv.append(t.tri_uv_leaf)
            tex_id.append(np.full(nf, len(info), np.int32))
            texels.append(tx)
            info.append((first, side))
            first += side * side
        self.has_textures = bool(info)
        self._texels_host = None
        self.t_tri_uv = torch.from_numpy(np.concatenate(uv)).to(dev) if self.has_textures else None
        self.t_tri_tex = torch.from_numpy(np.concatenate(tex_id)).to(dev) if self.has_textures else None
        self.t_texels = torch.from_numpy(np.concatenate(texels)).to(dev) if self.has_textures else None
        self.tex_info_host = np.asarray(info, np.int64).reshape(-1, 2)
        self.t_tex_info = torch.from_numpy(self.tex_info_host).to(dev) if self.has_textures else None
        self.n_texels = first

    def texture_inputs(self, mesh: int) -> dict:
        """The texture arrays of mesh ``mesh`` by original face id, with the
        model's whole texel table (the oracle's render_rgb inputs)."""
        t = self.topo[mesh]
        nf = t.mesh.num_tris
        out = {"tri_tex": np.full(nf, -1, np.int32), "tri_uv": np.zeros((nf, 3, 2), np.float32),
               "texels": np.zeros((1, 3), np.float32), "tex_info": np.zeros((1, 2), np.int64)}
        if not self.has_textures or t.tri_uv_leaf is None:
            return out
        base = int(self.mesh_off["f"][mesh])
        tid = self.t_tri_tex[base:base + nf].cpu().numpy()
        out["tri_tex"][t.tri_map] = tid
        out["tri_uv"][t.tri_map] = t.tri_uv_leaf.reshape(nf, 3, 2)
        if getattr(self, "_texels_host", None) is None:
            self._texels_host = self.t_texels[:, :3].cpu().numpy()
        out["texels"] = self._texels_host
        out["tex_info"] = self.tex_info_host
        return out

    def _build_envs(self):
        n = self.numEnvs
        desc = np.zeros((n, len(ENV_DESC_FIELDS)), dtype=np.int32)
        F = {f: i for i, f in enumerate(ENV_DESC_FIELDS)}
        vb = nb = le = lt = 0
        carry = self.carries_lambda
        self.env_vbase = np.zeros(n + 1, dtype=np.int64)
        self.env_nbase = np.zeros(n + 1, dtype=np.int64)
        for i in range(n):
            k = int(self.env_mesh[i])
            t = self.topo[k]
            m = t.mesh
            row = desc[i]
            row[F["vbase"]], row[F["nv"]] = vb, m.num_vertices
            row[F["ebase"]], row[F["ne"]] = self.mesh_off["e"][k], m.num_edges
            row[F["tbase"]], row[F["nt"]] = self.mesh_off["t"][k], m.num_tets
            row[F["fbase"]], row[F["nf"]] = self.mesh_off["f"][k], m.num_tris
            row[F["nodebase"]], row[F["nn"]] = self.mesh_off["n"][k], t.num_nodes
            row[F["boundbase"]] = nb
            row[F["lvlbase"]], row[F["nlvl"]] = self.mesh_off["l"][k], t.num_groups
            row[F["bvh_depth"]] = t.max_level + 1
            row[F["lam_ebase"]], row[F["lam_tbase"]] = (le, lt) if carry else (0, 0)
            row[F["dp_exp"]] = self._dp_exp(k)
            row[F["vtopo_base"]] = self.mesh_off["vt"][k]
            row[F["max_terms"]] = self._max_terms(k)
            vb += m.num_vertices
            nb += t.num_nodes
            if carry:
                le += m.num_edges
                lt += m.num_tets
            self.env_vbase[i + 1] = vb
            self.env_nbase[i + 1] = nb
        if vb >= 2 ** 31 or nb >= 2 ** 31:
            raise ValueError("state exceeds int32 indexing")
        self.env_desc_host = desc
        self.env_desc = torch.from_numpy(desc).to(self.device)
        self.numVertices = vb
        self.numNodes = nb
        self.num_lam_e, self.num_lam_t = le, lt
        self.max_nv = int(max(m.num_vertices for m in self.meshes)) if self.meshes else 0
        work = np.array([self.meshes[k].num_tets * 4 + self.meshes[k].num_edges for k in self.env_mesh])
        self.work = work

    def _dp_exp(self, k: int) -> int:
        """Exponent of the fixed-point quantum q of mesh k's Jacobi sums in
        the batched frames: q = 2^(floor(log2 D) - 26), D the mesh diagonal.
        A term is clamped to 2^24 - 1 quanta (D/8 .. D/4); its rounding error
        q/2 is 1/16 of the float32 spacing of a coordinate of size D.  The
        sums of a vertex's at most 128 terms stay inside int32 for any q."""
        diag = max(self.meshes[k].bbox_diagonal(), 1e-6)
        return int(math.floor(math.log2(diag))) - 26

    def _max_terms(self, k: int) -> int:
        """Most Jacobi terms one vertex of mesh k receives in one pass of the
        batched frame: a tet adds one term to each of its vertices (volume and
        Neo-Hookean summed), an edge one to each end; the collision pass adds
        one per incident surface triangle."""
        cache = self.__dict__.setdefault("_terms", {})
        if k not in cache:
            m = self.meshes[k]
            inc = np.zeros(m.num_vertices, np.int64)
            if self.vol_mode != OFF or self.nh_on:
                inc += np.bincount(m.tets.reshape(-1), minlength=m.num_vertices)
            if self.dist_mode != OFF:
                inc += np.bincount(m.edges.reshape(-1), minlength=m.num_vertices)
            terms = int(inc.max()) if inc.size else 0
            if self.collision and m.num_tris:
                terms = max(terms, int(np.bincount(m.tris.reshape(-1), minlength=m.num_vertices).max()))
            cache[k] = terms
        return cache[k]

    def _build_state(self, initial_velocity, tips, bases, radii):
        dev = self.device
        V = self.numVertices
        x0 = np.zeros((V, 4), dtype=np.float32)
        for i in range(self.numEnvs):
            m = self.meshes[int(self.env_mesh[i])]
            a, b = self.env_vbase[i], self.env_vbase[i + 1]
            x0[a:b, :3] = m.verts
            x0[a:b, 3] = m.inv_mass
        v0 = np.zeros((V, 4), dtype=np.float32)
        if initial_velocity is not None:
            v0[:, :3] = np.asarray(initial_velocity, dtype=np.float32).reshape(V, 3)
        self.initialVertex = torch.from_numpy(x0).to(dev)
        self.initialVelocity = torch.from_numpy(v0).to(dev)
        self.x = self.initialVertex.clone()
        self.p = self.initialVertex.clone()
        self.v = self.initialVelocity.clone()
        self.drag = torch.zeros(V, dtype=torch.float32, device=dev)
        self.lam_e = torch.zeros(max(self.num_lam_e, 1), dtype=torch.float32, device=dev) if self.carries_lambda else None
        self.lam_t = torch.zeros(max(self.num_lam_t, 1), dtype=torch.float32, device=dev) if self.carries_lambda else None
        self.bvh_lo = torch.zeros((max(self.numNodes, 1), 4), dtype=torch.float32, device=dev)
        self.bvh_hi = torch.zeros((max(self.numNodes, 1), 4), dtype=torch.float32, device=dev)
        t0 = np.zeros((self.numEnvs, 4, 3), dtype=np.float32) if tips is None else np.asarray(tips, np.float32)
        self.initialTips = torch.from_numpy(t0.reshape(self.numEnvs, 4, 3).copy()).to(dev)
        self.laparoscopeTip = self.initialTips.clone()
        # instrument capsules (laparoscope.py:375-386): segment base -> tip of
        # points 0..2 with a radius; point 3 is the drag point (radius 0)
        c0 = self.initialTips.cpu().numpy() if bases is None else np.asarray(bases, np.float32)
        self.initialBases = torch.from_numpy(c0.reshape(self.numEnvs, 4, 3).copy()).to(dev)
        self.laparoscopeBase = self.initialBases.clone()
        r0 = np.zeros((self.numEnvs, 4), np.float32) if radii is None else np.asarray(radii, np.float32)
        self.laparoscopeRadius = torch.from_numpy(r0.reshape(self.numEnvs, 4).copy()).to(dev)
        self.laparoscopeInCollision = torch.zeros(max(self.numEnvs, 1), dtype=torch.int32, device=dev)
        vc = np.concatenate([self.meshes[int(k)].colors for k in self.env_mesh]) if self.numEnvs else np.zeros((0, 3))
        self.vertexColors = torch.from_numpy(vc.astype(np.float32)).to(dev)
        self.normals = torch.zeros((V, 4), dtype=torch.float32, device=dev)
        self.sched = torch.zeros(2, dtype=torch.int32, device=dev)      # frame work counters
        self.status = torch.zeros(max(self.numEnvs, 1), dtype=torch.int32, device=dev)   # FFSRL_STATUS_* bits
        self.dp = None
        self.cnt = None

    def _make_batch(self) -> Batch:
        b = Batch()
        b.n_env = self.numEnvs
        b.max_nv = self.max_nv
        b.n_tet_total = self.n_tet_total
        b.envs = _ptr(self.env_desc)
        b.edges = _ptr(self.t_edges)
        b.edge_rest = _ptr(self.t_edge_rest)
        b.tets = _ptr(self.t_tets)
        b.tets16 = _ptr(self.t_tets16)
        b.edges16 = _ptr(self.t_edges16)
        b.tet_rest_volume = _ptr(self.t_tet_rest_volume)
        b.tet_inv_rest = _ptr(self.t_tet_inv_rest)
        b.tris = _ptr(self.t_tris)
        b.node_left = _ptr(self.t_node_left)
        b.node_first = _ptr(self.t_node_first)
        b.node_count = _ptr(self.t_node_count)
        b.level_nodes = _ptr(self.t_level_nodes)
        b.level_info = _ptr(self.t_level_info)
        b.level_off = _ptr(self.t_level_off)
        b.vf_off = _ptr(self.t_vf_off)
        b.vf_idx = _ptr(self.t_vf_idx)
        b.tri_uv = _ptr(self.t_tri_uv)
        b.tri_tex = _ptr(self.t_tri_tex)
        b.texels = _ptr(self.t_texels)
        b.tex_info = _ptr(self.t_tex_info)
        b.n_texels = self.n_texels
        b.x = _ptr(self.x)
        b.p = _ptr(self.p)
        b.v = _ptr(self.v)
        b.drag = _ptr(self.drag)
        b.lam_e = _ptr(self.lam_e)
        b.lam_t = _ptr(self.lam_t)
        b.bvh_lo = _ptr(self.bvh_lo)
        b.bvh_hi = _ptr(self.bvh_hi)
        b.tips = _ptr(self.laparoscopeTip)
        b.caps_base = _ptr(self.laparoscopeBase)
        b.caps_radius = _ptr(self.laparoscopeRadius)
        b.in_collision = _ptr(self.laparoscopeInCollision)
        b.sched = _ptr(self.sched)
        b.dp = _ptr(self.dp)
        b.cnt = _ptr(self.cnt)
        b.split_acc = _ptr(getattr(self, "split_acc", None))
        b.split_ctr = _ptr(getattr(self, "split_ctr", None))
        b.status = _ptr(self.status)
        fw = getattr(self, "fold_ws", None)
        b.fold_ws = _ptr(fw)
        b.fold_stride = self.numVertices if fw is not None else 0
        b.host_envs = self.env_desc_host.ctypes.data
        b.host_level_off = self.level_off_host.ctypes.data
        return b

    def _build_instrument(self, laps):
        """Transforms, heights, clamp state and workspace flags of every env's
        laparoscope (``modelDO.py:1208-1225``)."""
        dev = self.device
        xf = np.concatenate([lap.stored_transforms() for lap in laps])           # (n*7, 4, 4) transposed
        self.laparoscopeXForm = torch.from_numpy(xf).to(dev)
        self.laparoscopeInitialXForm = self.laparoscopeXForm.clone()
        h = np.stack([lap.initial_points()[3] for lap in laps]).astype(np.float32)
        self.laparoscopeHeight = torch.from_numpy(h).to(dev)
        n = self.numEnvs
        self.laparoscopeDragCutHeat = torch.zeros(3 * n, dtype=torch.int32, device=dev)
        self.laparoscopeDragCutHeatUpdate = torch.zeros(3 * n, dtype=torch.int32, device=dev)
        self.laparoscopeWorkspaceViolation = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self._workspace = torch.tensor([self.workspaceLow, self.workspaceHigh], dtype=torch.float32, device=dev)
        ins = Instrument()
        ins.n_env = n
        ins.xform = _ptr(self.laparoscopeXForm)
        ins.height = _ptr(self.laparoscopeHeight)
        ins.drag_cut_heat = _ptr(self.laparoscopeDragCutHeat)
        ins.drag_cut_heat_update = _ptr(self.laparoscopeDragCutHeatUpdate)
        ins.violation = _ptr(self.laparoscopeWorkspaceViolation)
        self._instrument = ins

    @property
    def instrument(self) -> Instrument:
        if self._instrument is None:
            raise RuntimeError("this model has no laparoscopes (pass laparoscopes= to SimModelDO)")
        return self._instrument

    def _actions(self, a, width: int) -> torch.Tensor:
        t = torch.as_tensor(a, dtype=torch.float32, device=self.device).reshape(-1)
        if t.numel() != width * self.numEnvs:
            raise ValueError(f"expected {width} action values per environment, got {t.numel()}")
        return t.contiguous()

    # ------------------------------------------------------------- instrument
    def applyActions(self, actions):
        """``modelDO.py:1331-1362``: per env (rod insertion, rod roll, rotation-
        centre pitch, yaw, jaw opening in radians); moves the instruments and
        reselects the grasp of envs whose clamp state changed."""
        a = self._actions(actions, 5)
        lib.ffsrl_instrument_actions(self._batch, self.instrument, a.data_ptr(),
                                     float(self.globalLaparoscopeDragLookupRadius), self.stream())

    def applyCartesianActions(self, cartesianActions):
        """``modelDO.py:1364-1383``: move each env's drag point by (dx, dy, dz)
        about its rotation centre."""
        a = self._actions(cartesianActions, 3)
        lib.ffsrl_instrument_cartesian(self._batch, self.instrument, a.data_ptr(), None, None, self.stream())

    def applyCartesianActionsInWorkspace(self, cartesianActions):
        """``modelDO.py:1385-1404``: as applyCartesianActions, dropping the
        components that would leave [workspaceLow, workspaceHigh]."""
        a = self._actions(cartesianActions, 3)
        w = self._workspace
        lib.ffsrl_instrument_cartesian(self._batch, self.instrument, a.data_ptr(), w[0].data_ptr(),
                                       w[1].data_ptr(), self.stream())

    def setRotationCenter(self, laparoscopeTranslation, laparoscopeRotation):
        """``modelDO.py:1570-1587``: place each env's rotation centre at
        T(translation) Rx Ry Rz(rotation)."""
        t = self._actions(laparoscopeTranslation, 3)
        r = self._actions(laparoscopeRotation, 3)
        lib.ffsrl_instrument_set_rcm(self._batch, self.instrument, t.data_ptr(), r.data_ptr(), self.stream())

    def setRotationCenterInitialPosition(self, laparoscopeTranslation, laparoscopeRotation):
        """``modelDO.py:1416-1419``."""
        self.setRotationCenter(laparoscopeTranslation, laparoscopeRotation)
        self.laparoscopeInitialXForm.copy_(self.laparoscopeXForm)

    def setEffectorInitialPosition(self, cartesianActions):
        """``modelDO.py:1406-1414``."""
        self.applyCartesianActions(cartesianActions)
        self.laparoscopeInitialXForm.copy_(self.laparoscopeXForm)
        self.initialBases.copy_(self.laparoscopeBase)
        self.initialTips.copy_(self.laparoscopeTip)

    def forceLaparoscopeClamp(self, envs):
        """``modelDO.py:1524-1531``: close the jaws of the envs with envs[i] > 0."""
        e = torch.clamp(torch.as_tensor(envs, dtype=torch.float32, device=self.device).reshape(-1), 0.0, 1.0)
        a = torch.zeros((self.numEnvs, 5), dtype=torch.float32, device=self.device)
        a[:, 4] = -e
        self.applyActions(a)

    def forceLaparoscopeClampVertex(self, vertexId: int, envs, on: float = 1.0, animate: bool = True):
        """``modelDO.py:1533-1544`` (``forceClampsDragKernel``, ``:282-289``):
        grasp weight ``on`` for local vertex ``vertexId`` of every env with
        envs[i] != 0.  ``animate`` is accepted and unused, as in the reference."""
        ids = torch.nonzero(torch.as_tensor(envs, device=self.device).reshape(-1)).flatten().cpu().numpy()
        for i in ids:
            nv = int(self.env_vbase[i + 1] - self.env_vbase[i])
            if not 0 <= int(vertexId) < nv:
                raise IndexError(f"vertex {vertexId} outside env {int(i)} ({nv} vertices)")
            self.drag[int(self.env_vbase[i]) + int(vertexId)] = float(on)

    def transformSimMeshData(self) -> None:
        """``modelDO.py:1441-1444`` copies simulation vertices into the render
        point arrays.  Here the render triangles index the simulation
        vertices directly, so the renderer always sees the current positions
        and there is nothing to copy."""

    def resetStochasticModel(self, laparoscopeTranslation, laparoscopeRotation):
        """``modelDO.py:1624-1645``: every env's positions and instrument back
        to their initial state (clamp state cleared), then the instrument
        re-aimed at the given rotation centre (``setRotationCenter``).  As in
        the reference, velocities, predicted positions and grasp weights are
        left as they are."""
        self.x.copy_(self.initialVertex)
        self.laparoscopeTip.copy_(self.initialTips)
        self.laparoscopeBase.copy_(self.initialBases)
        if self._instrument is not None:
            self.laparoscopeXForm.copy_(self.laparoscopeInitialXForm)
            self.laparoscopeDragCutHeat.zero_()
            self.laparoscopeDragCutHeatUpdate.zero_()
            self.setRotationCenter(laparoscopeTranslation, laparoscopeRotation)

    def getLaparoscopeWorkspaceViolation(self):
        """``modelDO.py:1508-1512``."""
        return self.laparoscopeWorkspaceViolation[:self.numEnvs]

    def set_bvh(self, mesh: int, bvh: dict) -> None:
        """Replace mesh ``mesh``'s BVH topology with ``bvh`` (keys ``left``,
        ``first``, ``count``, ``level``, ``tri_map``, ``num_nodes``,
        ``max_level``; ``SimBVH(load=True)`` reads them from a reference
        ``.npz``).  Every array that depends on the topology is rebuilt; the
        simulation state is kept; the bounds are zero until the next refit."""
        old = self.topo[mesh]
        validate_bvh(bvh, self.meshes[mesh].num_tris)
        self.topo[mesh] = MeshTopology(self.meshes[mesh], 8, old.constraint_order, old.chunk, bvh=bvh)
        self._build_topology()
        self._build_envs()
        dev = self.device
        self.bvh_lo = torch.zeros((max(self.numNodes, 1), 4), dtype=torch.float32, device=dev)
        self.bvh_hi = torch.zeros((max(self.numNodes, 1), 4), dtype=torch.float32, device=dev)
        self._batch = self._make_batch()
        self._order = None

    def enable_plain(self):
        """Allocate the dP / count scratch of the plain (per-kernel) form."""
        if self.dp is None:
            self.dp = torch.zeros((3, self.numVertices), dtype=torch.float32, device=self.device)
            self.cnt = torch.zeros(self.numVertices, dtype=torch.int32, device=self.device)
            self._batch.dp = _ptr(self.dp)
            self._batch.cnt = _ptr(self.cnt)

    def enable_shared(self, fold: Optional[bool] = None):
        """Allocate the global Jacobi sums and chunk counters of the shared-env
        frame (``ffsrl_frame_shared``); both start at zero and every frame
        leaves them at zero.  With ``fold`` (the default) also the workspace
        of its folded form (7 x 16 B per vertex), which envs of at most 5,104
        vertices take when the frame has no collision pass; ``fold=False``
        frees it, so every env takes the unfolded form (the results are the
        same, bit for bit); ``None`` keeps the current choice.  Probes the
        device's XCD count, so call it before capturing a shared frame into a
        graph."""
        if getattr(self, "split_acc", None) is None:
            self.split_acc = torch.zeros((self.numVertices, 2), dtype=torch.int64, device=self.device)
            self.split_ctr = torch.zeros((self.numEnvs, 4), dtype=torch.int32, device=self.device)
            self._batch.split_acc = _ptr(self.split_acc)
            self._batch.split_ctr = _ptr(self.split_ctr)
            if lib.ffsrl_xcd_count() <= 0:
                raise RuntimeError(f"ffsrl_xcd_count: {lib.ffsrl_last_error().decode()}")
        if fold is None:
            fold = getattr(self, "_fold", True)
        self._fold = bool(fold)
        if fold and getattr(self, "fold_ws", None) is None:
            self.fold_ws = torch.zeros((7, self.numVertices, 2), dtype=torch.int64, device=self.device)
        elif not fold:
            self.fold_ws = None
        self._batch.fold_ws = _ptr(self.fold_ws)
        self._batch.fold_stride = self.numVertices if self.fold_ws is not None else 0

    def enable_parts(self, tets_per_part: int = PART_TETS, parts: Optional[Sequence[int]] = None) -> Parts:
        """Partition every mesh's constraints for the partitioned frame
        (``ffsrl_frame_parts``) and allocate its sums and counters (zero, and
        left at zero by every frame).  A mesh gets ceil(tets / tets_per_part)
        parts (edges when the scheme has no tet constraints), more while a
        part has more vertices than one workgroup's LDS holds (an env in one
        part needs 32 B of LDS per vertex, a part of several 56 B); ``parts``
        gives the count per mesh instead.  The partition follows the current
        constraint order: ``set_constraint_order`` rebuilds it."""
        use_t = self.vol_mode != OFF or self.nh_on
        use_e = self.dist_mode != OFF
        cap = int(lib.ffsrl_parts_max_local())
        solo_cap = int(lib.ffsrl_parts_max_solo())

        def build(k):
            t = self.topo[k]
            m = t.mesh
            items = m.num_tets if use_t else m.num_edges
            P = int(parts[k]) if parts is not None else min(PART_MAX, max(1, -(-items // int(tets_per_part))))
            if parts is None and P == 1 and m.num_vertices > solo_cap:
                P = 2
            # a part's vertices grow with its share of the mesh: start near the LDS bound
            if parts is None and P > 1:
                P = max(P, min(PART_MAX, -(-m.num_vertices * 5 // (4 * cap))))
            while True:
                r = partition_mesh(t.tets, t.edges, m.num_vertices, P, use_t, use_e,
                                   tris=t.tris_leaf if self.collision else None)
                limit = solo_cap if P == 1 else cap
                if r["max_local"] <= limit or parts is not None or P >= PART_MAX:
                    break
                P += 1
            if r["max_local"] > (solo_cap if P == 1 else cap):
                raise ValueError(f"mesh {k}: a part of {r['max_local']} vertices exceeds the "
                                 f"{solo_cap if P == 1 else cap} that one workgroup holds ({P} parts)")
            return r

        built = _parallel(build, range(len(self.topo)))
        first = np.cumsum([0] + [r["parts"] for r in built])
        voff = np.cumsum([0] + [r["verts"].size for r in built])
        toff = np.cumsum([0] + [r["tris16"].size for r in built])
        rows = np.concatenate([r["rows"] for r in built])
        rows[:, 4] += np.repeat(voff[:-1], [r["parts"] for r in built])
        rows[:, 6] += np.repeat(toff[:-1], [r["parts"] for r in built])
        env_parts = np.stack([first[self.env_mesh], first[self.env_mesh + 1] - first[self.env_mesh]], axis=1)
        self._parts_host = np.ascontiguousarray(env_parts.astype(np.int32))
        dev = self.device
        cat = np.concatenate
        self.p_rows = torch.from_numpy(rows.astype(np.int32)).to(dev)
        self.p_env = torch.from_numpy(self._parts_host).to(dev)
        self.p_verts = torch.from_numpy(cat([r["verts"] for r in built])).to(dev)
        self.p_tets16 = torch.from_numpy(cat([r["tets16"] for r in built])).to(dev)
        self.p_edges16 = torch.from_numpy(cat([r["edges16"] for r in built])).to(dev)
        # with the collision pass, the parts' triangles (one entry at least, so the table is never NULL)
        self.p_tris16 = (torch.from_numpy(cat([r["tris16"] for r in built] + [np.zeros(1, np.int64)])).to(dev)
                         if self.collision else None)
        if getattr(self, "p_sums", None) is None:
            self.p_sums = torch.zeros((3, self.numVertices, 2), dtype=torch.int64, device=dev)
            self.p_ctr = torch.zeros(4 * self.numEnvs + 64, dtype=torch.int32, device=dev)
            if lib.ffsrl_xcd_count() <= 0:
                raise RuntimeError(f"ffsrl_xcd_count: {lib.ffsrl_last_error().decode()}")
        pt = Parts()
        pt.n_parts = int(rows.shape[0])
        pt.max_local = max([r["max_local"] for r in built if r["parts"] > 1], default=0)
        pt.max_solo = max([r["max_local"] for r in built if r["parts"] == 1], default=0)
        pt.parts, pt.env_parts, pt.verts = _ptr(self.p_rows), _ptr(self.p_env), _ptr(self.p_verts)
        pt.tets16, pt.edges16, pt.tris16 = _ptr(self.p_tets16), _ptr(self.p_edges16), _ptr(self.p_tris16)
        pt.sums, pt.stride, pt.ctr = _ptr(self.p_sums), self.numVertices, _ptr(self.p_ctr)
        n_claims = int(self._parts_host[:, 1].sum())
        if getattr(self, "p_claims", None) is None or self.p_claims.numel() < n_claims:
            self.p_claims = torch.zeros(n_claims, dtype=torch.int32, device=dev)
        pt.claims, pt.n_claims = _ptr(self.p_claims), int(self.p_claims.numel())
        pt.host_env_parts = self._parts_host.ctypes.data
        self._parts = pt
        self._parts_cfg = dict(tets_per_part=tets_per_part, parts=parts)
        self.parts_per_mesh = [r["parts"] for r in built]
        return pt

    @property
    def parts(self) -> Parts:
        """The partition of ``enable_parts`` (built with its defaults on first use)."""
        if getattr(self, "_parts", None) is None:
            self.enable_parts()
        return self._parts

    def frame_status(self) -> torch.Tensor:
        """Per-env status bits of the batched and shared frames, sticky until
        ``clear_status``: STATUS_SATURATED (a fixed-point Jacobi term left
        its range or was not finite, so the frame no longer equals the plain
        form) and STATUS_INCOMPLETE (a shared frame could not complete)."""
        return self.status[:self.numEnvs]

    shared_status = frame_status

    def clear_status(self) -> None:
        self.status.zero_()

    def check_status(self) -> None:
        """Raises FFSRLError naming the envs with a nonzero status (a host
        synchronisation: call it at the caller's own sync points)."""
        st = self.frame_status().cpu().numpy()
        bad = np.nonzero(st)[0]
        if bad.size:
            kinds = []
            if (st & STATUS_SATURATED).any():
                kinds.append(f"{int(((st & STATUS_SATURATED) != 0).sum())} with a saturated fixed-point term")
            if (st & STATUS_INCOMPLETE).any():
                kinds.append(f"{int(((st & STATUS_INCOMPLETE) != 0).sum())} with an incomplete shared frame")
            raise FFSRLError(f"frame status set for {bad.size} envs ({', '.join(kinds)}), first env {int(bad[0])}")

    @property
    def batch(self) -> Batch:
        return self._batch

    def set_constraint_order(self, constraint_order: str = "local", chunk: int = LOCAL_CHUNK) -> None:
        """Re-order every mesh's constraints on the device in place (the
        autotuner's chunk knob).  Carried XPBD lambdas follow their
        constraints; positions are unaffected."""
        dev = self.device
        old = [(t.tet_perm, t.edge_perm) for t in self.topo]
        _parallel(lambda t: t.set_order(constraint_order, chunk), self.topo)
        cat = np.concatenate
        self.t_tets.copy_(torch.from_numpy(cat([t.tets for t in self.topo])))
        self.t_tets16.copy_(torch.from_numpy(_pack16(cat([t.tets for t in self.topo]))))
        self.t_edges16.copy_(torch.from_numpy(_pack16(cat([t.edges for t in self.topo]))))
        self.t_tet_rest_volume.copy_(torch.from_numpy(cat([t.tet_rest_volume for t in self.topo])))
        inv_rest = cat([t.tet_inv_rest for t in self.topo]).astype(np.float32)
        self.t_tet_inv_rest.copy_(torch.from_numpy(np.ascontiguousarray(inv_rest.T)))
        self.t_edges.copy_(torch.from_numpy(cat([t.edges for t in self.topo])))
        self.t_edge_rest.copy_(torch.from_numpy(cat([t.edge_rest for t in self.topo])))
        if self.lam_t is not None:
            for k, t in enumerate(self.topo):
                envs = np.nonzero(self.env_mesh == k)[0]
                for lam, base_f, o, new in ((self.lam_t, "lam_tbase", old[k][0], t.tet_perm),
                                            (self.lam_e, "lam_ebase", old[k][1], t.edge_perm)):
                    if o.size == 0 or envs.size == 0:
                        continue
                    inv_new = np.empty_like(new)
                    inv_new[new] = np.arange(new.size)
                    dst = torch.from_numpy(inv_new[o]).to(dev)           # old position -> new position
                    base = torch.from_numpy(self.env_desc_host[envs, ENV_DESC_FIELDS.index(base_f)].astype(np.int64))
                    base = base.to(dev)[:, None]
                    src = (base + torch.arange(o.size, device=dev)[None]).reshape(-1)
                    lam[(base + dst[None]).reshape(-1)] = lam[src].clone()
        self.constraint_order = constraint_order
        if getattr(self, "_parts", None) is not None:
            self.enable_parts(**self._parts_cfg)   # the partition follows the constraint order

    def dispatch_order(self) -> torch.Tensor:
        """Environments largest first: the hardware dispatcher then assigns
        them to CUs in longest-processing-time order."""
        if self._order is None:
            order = np.argsort(-self.work, kind="stable").astype(np.int32)
            self._order = torch.from_numpy(order).to(self.device)
        return self._order

    @staticmethod
    def stream() -> int:
        return torch.cuda.current_stream().cuda_stream

    # ---------------------------------------------------------------- getters
    def env_slice(self, i: int) -> slice:
        return slice(int(self.env_vbase[i]), int(self.env_vbase[i + 1]))

    def vertex(self) -> torch.Tensor:
        """(V, 3) view of all positions (``SimModelDO.vertex``)."""
        return self.x[:, :3]

    def env_vertices(self, i: int) -> torch.Tensor:
        return self.x[self.env_slice(i), :3]

    def getVertexPositionsTensor(self, vertexId):
        """``modelDO.py:1483``: positions of global vertex ids."""
        return self.x[:, :3][vertexId]

    def getVertexPositionsTensorMultiEnvs(self, vertexIds):
        """``modelDO.py:1489``: env-local vertex ids, one per env."""
        ids = torch.as_tensor(vertexIds, device=self.device).reshape(-1)
        base = torch.as_tensor(self.env_vbase[:-1], device=self.device)
        return self.x[:, :3][ids + base[: ids.numel()]]

    def getLaparoscopePositionsTensor(self):
        """``modelDO.py:1496``: drag point (index 3) of each env's instrument."""
        return self.laparoscopeTip[:, 3, :]

    def getLaparoscopeClampVertex(self, vertexId):
        return self.drag[vertexId]

    def getLaparoscopeInCollision(self):
        """``modelDO.py:1513-1517``: per env, the surface triangles found in
        contact, summed over every collision pass since the last reset."""
        return self.laparoscopeInCollision[:self.numEnvs]

    def resetCollisionInfo(self):
        """``modelDO.py:1519-1521`` (cleared in place: the batch keeps its pointer)."""
        self.laparoscopeInCollision.zero_()

    # ------------------------------------------------------------------ reset
    def resetFixedModel(self, envNums):
        """``modelDO.py:1589-1622``: restore positions, zero predicted positions,
        velocities and grasp weights, restore the instrument points."""
        for i in envNums:
            s = self.env_slice(int(i))
            self.x[s] = self.initialVertex[s]
            self.p[s] = 0.0
            self.p[s, 3] = self.initialVertex[s, 3]
            self.v[s] = 0.0
            self.drag[s] = 0.0
            self.laparoscopeTip[int(i)] = self.initialTips[int(i)]
            self.laparoscopeBase[int(i)] = self.initialBases[int(i)]
            if self._instrument is not None:
                k = int(i)
                self.laparoscopeXForm[7 * k:7 * k + 7] = self.laparoscopeInitialXForm[7 * k:7 * k + 7]
                self.laparoscopeDragCutHeat[3 * k:3 * k + 3] = 0
                self.laparoscopeDragCutHeatUpdate[3 * k:3 * k + 3] = 0

    def set_grasp(self, env: int, local_ids, weight: float = 1.0):
        """Force grasp weights on env-local vertices (``forceClampsDragKernel``,
        ``modelDO.py:281-289``)."""
        ids = torch.as_tensor(np.asarray(local_ids, dtype=np.int64) + int(self.env_vbase[env]), device=self.device)
        self.drag[ids] = weight
