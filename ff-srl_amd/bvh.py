"""SimBVH: per-environment triangle BVHs (``FF_SRL/FF_SRL/bvh.py:146-518``).

The topology of each mesh is built on the host by the library's binned-SAH
builder when the model is created (``SimModelDO`` holds it because the
batched frame kernel refits in the same launch); this class exposes the
reference's refit call, the node arrays and the ``.npz`` checkpoint format
(``save=True`` writes it, ``load=True`` replaces the built topology with a
saved one, as ``loadBVH`` does, ``bvh.py:176-197``).
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import lib
from .model import SimModelDO


class SimBVH:
    def __init__(self, device, simModel: SimModelDO, binsNumber: int = 8, save: bool = False, load: bool = False,
                 path: str = "bvh.npz"):
        if load:
            # the reference shares one topology among its envs: the saved one
            # goes to every mesh with the saved mesh's triangle count
            topo = self.topology_from_npz(self.load_topology(path))
            nf = len(topo["tri_map"])
            meshes = [k for k, m in enumerate(simModel.meshes) if m.num_tris == nf]
            if not meshes:
                raise ValueError(f"{path}: no mesh of the model has {nf} triangles")
            for k in meshes:
                simModel.set_bvh(k, topo)
        self.device = device
        self.simModel = simModel
        self.binsNumber = binsNumber
        self.numEnvs = simModel.numEnvs
        self.refitGraph = None
        self.refitBVH()
        if save:
            self.saveBVH(path)

    # Reference accessors (env 0 / mesh 0 topology, as the reference shares one).
    @property
    def numNodes(self) -> int:
        return self.simModel.topo[int(self.simModel.env_mesh[0])].num_nodes

    @property
    def maxLevel(self) -> int:
        return max(t.max_level for t in self.simModel.topo)

    @property
    def aabbMin(self) -> torch.Tensor:
        return self.simModel.bvh_lo[:, :3]

    @property
    def aabbMax(self) -> torch.Tensor:
        return self.simModel.bvh_hi[:, :3]

    def env_bounds(self, env: int):
        s = slice(int(self.simModel.env_nbase[env]), int(self.simModel.env_nbase[env + 1]))
        return self.simModel.bvh_lo[s, :3], self.simModel.bvh_hi[s, :3]

    def refitBVH(self, useGraph: bool = False, plain: bool = False) -> None:
        """Bottom-up refit of every env (``bvh.py:442-451``)."""
        m = self.simModel
        if plain:
            lib.ffsrl_bvh_refit_plain(m.batch, 0, m.numEnvs, m.stream())
        else:
            lib.ffsrl_bvh_refit(m.batch, 0, m.numEnvs, m.stream())

    def saveBVH(self, path: str, mesh: int = 0) -> None:
        """Reference ``.npz`` layout (``bvh.py:165-174``): arr_0..arr_6 =
        triId, aabbMin, aabbMax, leftNode, firstTriId, triCount, level of one
        mesh (bounds of its first env)."""
        t = self.simModel.topo[mesh]
        env = int(np.nonzero(self.simModel.env_mesh == mesh)[0][0])
        lo, hi = self.env_bounds(env)
        np.savez(path, t.tri_map, lo.cpu().numpy(), hi.cpu().numpy(), t.left, t.first, t.count, t.level)

    def printBVH(self, mesh: int = 0) -> None:
        """Text picture of mesh ``mesh``'s tree as ``printBVH`` draws it
        (``bvh.py:199-228``): one row per level, each node's first triangle
        at its column (children at the parent's column -+ 2^(h-2)), blanks
        elsewhere."""
        t = self.simModel.topo[mesh]
        height = t.max_level + 1
        width = 2 ** height
        grid = [[-1] * width for _ in range(height)]
        stack = [(0, width // 2, 0, height)]
        while stack:
            node, col, row, h = stack.pop()
            if h < 1:
                continue
            grid[row][col] = int(t.first[node])
            if t.left[node] > -1:
                step = 2 ** (h - 2) if h >= 2 else 0
                stack.append((int(t.left[node]) + 1, col + step, row + 1, h - 1))
                stack.append((int(t.left[node]), col - step, row + 1, h - 1))
        for r in grid:
            print(" ".join(" " if v == -1 else str(v) for v in r))

    @staticmethod
    def topology_from_npz(arrays: dict) -> dict:
        """The topology of a reference ``.npz`` BVH in the form ``MeshTopology``
        takes: leaves get left child -1 (the reference leaves the constructor's
        -1 there too, and marks an unsplit root with 0, ``bvh.py:108,480``)."""
        count = np.asarray(arrays["triCount"], np.int32)
        left = np.where(count > 0, -1, np.asarray(arrays["leftNode"], np.int32)).astype(np.int32)
        level = np.asarray(arrays["level"], np.int32)
        return {"left": left, "first": np.asarray(arrays["firstTriId"], np.int32), "count": count, "level": level,
                "tri_map": np.asarray(arrays["triId"], np.int32), "num_nodes": int(count.size),
                "max_level": int(level.max()) if level.size else 0}

    @staticmethod
    def load_topology(path: str):
        """Read a reference ``.npz`` BVH (``bvh.py:176-197``) without unpickling."""
        with np.load(path, allow_pickle=False) as z:
            return {k: z[f"arr_{i}"] for i, k in enumerate(
                ("triId", "aabbMin", "aabbMax", "leftNode", "firstTriId", "triCount", "level"))}
