"""SimIntegratorDO: one frame of every environment.

``stepModel`` mirrors ``SimIntegratorDO.stepModel`` (``FF_SRL/FF_SRL/
integrator.py:1640-1744``): S substeps of predict, I iterations of the
constraint projection with Jacobi averaging, the grasp pull and (with
``collision``) the instrument collision, the ground clamp and the integrate.
Three launch forms run the same arithmetic:

* batched (default): ``ffsrl_frame``, one launch for all environments and
  the whole substep loop (and optionally the BVH refit);
* shared (``shared=True``): ``ffsrl_frame_shared``, the same launch with all
  CUs sharing the environments (each env's frame cut into chunks that the
  workgroups of one XCD take in turn), for batches with fewer environments
  than CUs; bit-identical to the batched form;
* partitioned (``parts=True``): ``ffsrl_frame_parts``, each env's constraints
  cut into parts with one workgroup per part for the whole frame (state in
  LDS, only shared vertices exchanged per pass); bit-identical to the
  batched form, and it also runs envs above the batched frame's vertex
  limit;
* plain: the reference's launch model with one launch per kernel per
  environment (``integrator.py:1779-1929`` is the reference's per-env
  precedent), used as the comparison baseline.
"""
from __future__ import annotations

import torch

from ._lib import lib
from .model import SimModelDO


class SimIntegratorDO:
    def __init__(self, device="cuda:0", block: int = 1024, dispatch_sorted: bool = True, shared: bool = False,
                 parts: bool = False):
        self.device = device
        self.block = block
        self.dispatch_sorted = dispatch_sorted
        self.shared = shared
        self.parts = parts

    def configure(self, cfg: dict) -> None:
        """Launch shape from the autotuner: {block, dispatch ("lpt"/"table"), shared, ...}."""
        self.block = int(cfg["block"])
        self.dispatch_sorted = cfg["dispatch"] == "lpt"
        self.shared = bool(cfg.get("shared", False))
        self.parts = bool(cfg.get("parts", False))

    def autotune(self, simModel: SimModelDO, cache_path=None, frames: int = 2) -> dict:
        """Tune (or look up) the launch shape for this batch and use it."""
        from .autotune import Autotuner
        cfg = Autotuner(cache_path).tune(simModel, frames=frames)
        self.configure(cfg)
        return cfg

    def stepModel(self, simModel: SimModelDO, refit: bool = False) -> None:
        """One frame of all environments in one launch (``ffsrl_frame``, or
        ``ffsrl_frame_shared`` / ``ffsrl_frame_parts`` when the integrator is
        shared / partitioned)."""
        m = simModel
        order = m.dispatch_order().data_ptr() if self.dispatch_sorted else None
        if self.parts:
            lib.ffsrl_frame_parts(m.batch, m.params, m.parts, order, m.numEnvs, int(bool(refit)), m.stream())
        elif self.shared:
            m.enable_shared()
            lib.ffsrl_frame_shared(m.batch, m.params, order, m.numEnvs, self.block, int(bool(refit)), m.stream())
        else:
            lib.ffsrl_frame(m.batch, m.params, order, self.block, int(bool(refit)), m.stream())

    def stepModelDO(self, simModel: SimModelDO, refit: bool = False) -> None:
        """One frame in the reference's launch model: every kernel launched once
        per substep and iteration over all environments (``ffsrl_frame_do``;
        the plain form's kernels and float atomics), and with ``refit`` one
        refit launch per BVH refit group over all environments."""
        m = simModel
        m.enable_plain()
        lib.ffsrl_frame_do(m.batch, m.params, 0, m.numEnvs, int(bool(refit)), m.stream())

    def stepModelPlain(self, simModel: SimModelDO, env_begin: int = 0, env_end: int = None) -> None:
        """One frame with one launch per kernel per environment (the library
        loops over the environments of each call)."""
        m = simModel
        m.enable_plain()
        eb, ee = env_begin, m.numEnvs if env_end is None else env_end
        b, prm, s = m.batch, m.params, m.stream()
        reset = m.carries_lambda and (not m.lambda_legacy or m.simConstraints > 1)
        for _ in range(m.simSubsteps):
            lib.ffsrl_predict(b, prm, eb, ee, s)
            if reset:
                lib.ffsrl_zero_lambdas(b, eb, ee, s)
            for _ in range(m.simConstraints):
                lib.ffsrl_project_volume(b, prm, eb, ee, s)
                lib.ffsrl_project_distance(b, prm, eb, ee, s)
                lib.ffsrl_project_neohookean(b, prm, eb, ee, s)
                lib.ffsrl_apply(b, eb, ee, s)
                lib.ffsrl_grasp(b, prm, eb, ee, s)
                if m.collision:
                    lib.ffsrl_collide(b, prm, eb, ee, s)
            lib.ffsrl_bounds_integrate(b, prm, eb, ee, s)
