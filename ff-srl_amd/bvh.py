"""SimBVH: per-environment triangle BVHs (``FF_SRL/FF_SRL/bvh.py:146-518``).

The topology of each mesh is built on the host by the library's binned-SAH
builder when the model is created (``SimModelDO`` holds it because the
batched frame kernel refits in the same launch); this class exposes the
reference's refit call, the node arrays and the ``.npz`` checkpoint format.
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
            raise NotImplementedError("the topology is built with the model; load .npz through SimBVH.load_topology")
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

    @staticmethod
    def load_topology(path: str):
        """Read a reference ``.npz`` BVH (``bvh.py:176-197``) without unpickling."""
        with np.load(path, allow_pickle=False) as z:
            return {k: z[f"arr_{i}"] for i, k in enumerate(
                ("triId", "aabbMin", "aabbMax", "leftNode", "firstTriId", "triCount", "level"))}
