"""Environment surface for RL loops over the simulation path.

It follows the call pattern of the reference's demo harness
(``tests/testLiverRetraction.py:19-97``, ``TestStandalone``): a model, an
integrator, a BVH and a renderer built together;
``step(actions, actionsCartesian)`` clears the collision counters
(``:61``), applies the joint actions and then the Cartesian actions within
the workspace (``:63-71``), and steps every environment (``:73-84``), with
the frame captured once into a HIP graph and replayed when ``useGraph`` is
set; ``render()`` renders every environment (``:86-93``).  The demo's main
loop refits the BVH and renders after each step (``:147-150``); here the
refit runs inside the frame kernel.

Added for RL loops: ``reset(env_ids)`` (``resetFixedModel``,
``modelDO.py:1589``, as the demo's "r" key does) and ``observe()``, a
dictionary of torch tensors on the device (image, tool points, clamp state,
instrument contacts).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import torch

from ._lib import lib
from .bvh import SimBVH
from .integrator import SimIntegratorDO
from .model import SimModelDO
from .render import RaycastRendererDO, RenderMode


class SimEnvironment:
    """``numEnvs`` environments of one model stepped and rendered together.

    The frame (physics of every environment and the BVH refit) is one
    ``ffsrl_frame`` launch; with ``useGraph`` it is captured into a HIP graph
    on the first step and replayed afterwards (actions are applied before the
    replay, outside the graph, since they take the caller's tensors).  Results
    are the same with and without the graph.  A graph holds the array
    pointers, parameters and launch shape it was captured with; when any of
    them has changed since (a new constraint order, an autotuned launch shape,
    the shared frame switched on), the next step captures the frame again.
    """

    def __init__(self, simModel: SimModelDO, renderer: Optional[RaycastRendererDO] = None,
                 integrator: Optional[SimIntegratorDO] = None, simBVH: Optional[SimBVH] = None,
                 useGraph: bool = False):
        self.simModel = simModel
        self.device = simModel.device
        self.numEnvs = simModel.numEnvs
        self.simIntegrator = integrator if integrator is not None else SimIntegratorDO(self.device)
        self.simBVH = simBVH if simBVH is not None else SimBVH(self.device, simModel)
        self.renderer = renderer
        self.useGraph = bool(useGraph)
        self.capturedGraph = None
        self._graph_stream = None
        self._graph_key = None
        self.captures = 0  # frame graphs captured so far
        self.image = None
        self.frames = 0

    @classmethod
    def from_usd(cls, path: str, numEnvs: int, device="cuda:0", *, resolution: int = 64,
                 cameraPos=(0.0, 50.0, 50.0), lookAt=(0.0, 10.0, 0.0), lightPos=(0.0, 35.0, 50.0),
                 lightIntensity: float = 0.03, render_mode: int = RenderMode.depth_map, useGraph: bool = False,
                 bvh_path: Optional[str] = None, **model_kwargs) -> "SimEnvironment":
        """The demo's set-up (``testLiverRetraction.py:44-49``): the scene (a
        USD crate file or its converted ``.npz``) in every env, the BVH (from
        the reference's ``.npz`` cache when ``bvh_path`` is given, else built),
        and one camera looking at ``lookAt``."""
        model = SimModelDO.from_usd(path, numEnvs, device=device, **model_kwargs)
        bvh = SimBVH(device, model, load=bvh_path is not None, path=bvh_path or "bvh.npz")
        ren = RaycastRendererDO(device, model, cameraPos=cameraPos, lightPos=lightPos, lightIntensity=lightIntensity,
                                resolution=resolution, render_mode=render_mode)
        ren.lookAt(lookAt)
        return cls(model, renderer=ren, simBVH=bvh, useGraph=useGraph)

    # ------------------------------------------------------------------ step
    def _frame(self) -> None:
        m = self.simModel
        if not self.useGraph:
            self.simIntegrator.stepModel(m, refit=True)
            return
        cur = torch.cuda.current_stream(self.device)
        integ = self.simIntegrator
        if integ.shared:
            m.enable_shared()  # allocates and probes once: never inside the capture
        key = (bytes(m.batch), bytes(m.params), integ.block, integ.dispatch_sorted, integ.shared,
               m.dispatch_order().data_ptr())
        if self.capturedGraph is not None and key != self._graph_key:
            self.close()
        if self.capturedGraph is None:
            self._graph_stream = torch.cuda.Stream(self.device)
            self._graph_stream.wait_stream(cur)
            g = C.c_void_p()
            s = self._graph_stream.cuda_stream
            with torch.cuda.stream(self._graph_stream):
                lib.ffsrl_graph_begin(s)
                try:
                    integ.stepModel(m, refit=True)
                except BaseException:
                    # leave capture mode and drop the partial graph before re-raising
                    lib.ffsrl_graph_abort(s)
                    raise
                lib.ffsrl_graph_end(s, C.byref(g))
            self._graph_key = key
            self.capturedGraph = g
            self.captures += 1
        self._graph_stream.wait_stream(cur)
        lib.ffsrl_graph_launch(self.capturedGraph, self._graph_stream.cuda_stream)
        cur.wait_stream(self._graph_stream)

    def step(self, actions=None, actionsCartesian=None, render: bool = False) -> dict:
        """One frame of every environment after the actions: ``actions`` are
        5 joint values per env (``applyActions``), ``actionsCartesian`` 3
        drag-point displacements per env applied within the workspace
        (``applyCartesianActionsInWorkspace``).  Returns ``observe()``; with
        ``render`` the image is rendered first."""
        m = self.simModel
        m.resetCollisionInfo()
        if actions is not None:
            m.applyActions(actions)
        if actionsCartesian is not None:
            m.applyCartesianActionsInWorkspace(actionsCartesian)
        self._frame()
        self.frames += 1
        if render:
            self.render()
        return self.observe()

    def render(self) -> torch.Tensor:
        """``renderNew`` of every environment: ``(numEnvs, H, W, 3)``."""
        if self.renderer is None:
            raise RuntimeError("this environment has no renderer")
        self.image = self.renderer.renderNew(self.simBVH, self.simModel)
        return self.image

    # ----------------------------------------------------------------- reset
    def reset(self, env_ids: Optional[Sequence[int]] = None, render: bool = False) -> dict:
        """Restore the listed environments (all when None) to their initial
        state (``resetFixedModel``), refit their BVHs and return
        ``observe()``."""
        ids = range(self.numEnvs) if env_ids is None else [int(i) for i in env_ids]
        for i in ids:
            if not 0 <= i < self.numEnvs:
                raise IndexError(f"environment {i} of {self.numEnvs}")
        m = self.simModel
        m.resetFixedModel(ids)
        m.resetCollisionInfo()
        self.simBVH.refitBVH()
        if render:
            self.render()
        return self.observe()

    def observe(self) -> dict:
        """Device tensors of the current state: ``tool`` (numEnvs, 4, 3)
        instrument points (clamp tips and drag point), ``clamp`` (numEnvs, 3)
        clamp / grasp state words, ``inCollision`` (numEnvs,) contact counts
        of the last step (instrument collision only), and ``image``, the last
        render (None before the first)."""
        m = self.simModel
        obs = {
            "tool": m.laparoscopeTip.reshape(self.numEnvs, 4, 3),
            "inCollision": m.laparoscopeInCollision,
            "image": self.image,
        }
        if getattr(m, "laparoscopeDragCutHeat", None) is not None:
            obs["clamp"] = m.laparoscopeDragCutHeat.reshape(self.numEnvs, 3)
        return obs

    def close(self) -> None:
        if self.capturedGraph is not None:
            torch.cuda.synchronize(self.device)
            lib.ffsrl_graph_destroy(self.capturedGraph)
            self.capturedGraph = None
            self._graph_key = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
