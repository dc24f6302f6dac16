"""Autotuner of the batched frame: picks the launch shape for a batch of
environments by timing the candidates on the device, and caches the choice
keyed on the env count and the tets per env.

Knobs (all leave the results unchanged; the frame's Jacobi sums are integers):
  block     threads per environment workgroup (256, 512, 1024);
  dispatch  environment dispatch order: "lpt" (largest first) or "table";
  chunk     run length of the colour sort inside the Morton constraint order
            (the lanes of one LDS atomic rarely share a vertex within a run,
            and a wavefront's position gathers stay local);
  shared    the shared-env frame (every CU on the batch's envs, each env's
            frame cut into chunks), tried when there are at most half as
            many envs as CUs;
  parts     the partitioned frame (each env's constraints cut into parts,
            one 1024-thread workgroup per part with the part's state in
            LDS): 0 (not partitioned), or the tets per part (PART_WHOLE:
            as few parts as a workgroup's LDS allows).

A trial runs whole frames on a snapshot of the state, which is restored
afterwards, so tuning does not advance the simulation.  The cache is a JSON
file; its key is (env-count bucket, mean-tets bucket, max-tets bucket,
constraint set), buckets being powers of two.
"""
from __future__ import annotations

import itertools
import json
import math
import os
from typing import Dict, Optional

import numpy as np
import torch

from ._lib import FFSRLError, lib
from .model import LOCAL_CHUNK, PART_TETS, SimModelDO

BLOCKS = (256, 512, 1024)
DISPATCH = ("lpt", "table")
CHUNKS = (64, 128, 256)
PART_WHOLE = 1 << 30
PART_SIZES = (PART_WHOLE, PART_TETS)
DEFAULT = {"block": 1024, "dispatch": "lpt", "chunk": LOCAL_CHUNK, "shared": False, "parts": 0}


def parts_candidates(model: SimModelDO) -> tuple:
    """(0,) or (0, *PART_SIZES): the partitioned frame (tets per part; 0 =
    not partitioned) is a candidate where the device has an XCD map."""
    if lib.ffsrl_xcd_count() <= 0:
        return (0,)
    return (0,) + PART_SIZES


def shared_candidates(model: SimModelDO) -> tuple:
    """(False,) or (False, True): the shared-env frame is a candidate when the
    batch has at most half as many envs as the device has CUs."""
    cus = torch.cuda.get_device_properties(model.device).multi_processor_count
    if 2 * model.numEnvs > cus or lib.ffsrl_xcd_count() <= 0:   # no XCD map: the shared frame is unavailable
        return (False,)
    return (False, True)


def _bucket(x: float) -> int:
    return 1 << max(0, int(math.ceil(math.log2(max(x, 1.0)))))


def key_of(model: SimModelDO) -> str:
    """Cache key: env-count and tets buckets, the constraint set, the
    collision pass (a different kernel) and whether the shared-env frame can
    run on this device (an XCD map)."""
    tets = np.array([model.meshes[k].num_tets for k in model.env_mesh], dtype=np.float64)
    flags = (f"d{model.dist_mode}v{model.vol_mode}n{int(model.nh_on)}s{model.simSubsteps}i{model.simConstraints}"
             f"c{int(model.collision)}x{int(lib.ffsrl_xcd_count() > 0)}k2")
    return f"envs{_bucket(model.numEnvs)}-tets{_bucket(tets.mean())}-max{_bucket(tets.max())}-{flags}"


class Autotuner:
    def __init__(self, cache_path: Optional[str] = None):
        self.cache_path = cache_path
        self.cache: Dict[str, dict] = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def save(self) -> None:
        if self.cache_path:
            os.makedirs(os.path.dirname(os.path.abspath(self.cache_path)), exist_ok=True)
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.cache_path)

    @staticmethod
    def _state(model: SimModelDO):
        # the contact counters and status words too: trial frames add to them
        return [model.x, model.p, model.v, model.bvh_lo, model.bvh_hi, model.laparoscopeInCollision, model.status]

    @staticmethod
    def _lambdas(model: SimModelDO):
        return [] if model.lam_t is None else [model.lam_t, model.lam_e]

    @staticmethod
    def _time(model: SimModelDO, cfg: dict, frames: int, stream) -> float:
        order = model.dispatch_order().data_ptr() if cfg["dispatch"] == "lpt" else None
        s = stream.cuda_stream
        if cfg.get("parts"):
            if getattr(model, "_parts_cfg", {}).get("tets_per_part") != cfg["parts"]:
                model.enable_parts(tets_per_part=cfg["parts"])
            run = lambda: lib.ffsrl_frame_parts(model.batch, model.params, model.parts, order, model.numEnvs, 1, s)
        elif cfg.get("shared"):
            model.enable_shared()
            run = lambda: lib.ffsrl_frame_shared(model.batch, model.params, order, model.numEnvs, cfg["block"], 1, s)
        else:
            run = lambda: lib.ffsrl_frame(model.batch, model.params, order, cfg["block"], 1, s)
        run()                                                                           # warm
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(frames):
            run()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / frames

    def tune(self, model: SimModelDO, frames: int = 2, force: bool = False) -> dict:
        """Best {block, dispatch, chunk, shared, parts} for this batch (cached),
        applied to the model's constraint order (and partition); the state is
        left as it was."""
        key = key_of(model)
        if not force and key in self.cache:
            cfg = dict(self.cache[key]["best"])
            self.apply(model, cfg)
            return cfg
        stream = torch.cuda.current_stream(model.device)
        orig = (model.topo[0].constraint_order, model.topo[0].chunk) if model.topo else ("local", LOCAL_CHUNK)
        state = [(t, t.clone()) for t in self._state(model)]
        lam0 = [(t, t.clone()) for t in self._lambdas(model)]
        results = []
        for chunk in CHUNKS:
            model.set_constraint_order("local", chunk)          # carried lambdas follow their constraints
            lam = [(t, t.clone()) for t in self._lambdas(model)]
            torch.cuda.synchronize(model.device)
            for block, dispatch, shared, parts in itertools.product(BLOCKS, DISPATCH, shared_candidates(model),
                                                                    parts_candidates(model)):
                if parts and (block != 1024 or shared):   # the partitioned frame runs 1024-thread parts
                    continue
                cfg = {"block": block, "dispatch": dispatch, "chunk": chunk, "shared": shared, "parts": parts}
                try:
                    ms = self._time(model, cfg, frames, stream)
                except (FFSRLError, ValueError):          # a mesh the partition cannot hold
                    ms = float("inf")
                for t, saved in state + lam:
                    t.copy_(saved)
                results.append((ms, cfg))
        model.set_constraint_order(*orig)
        for t, saved in lam0:
            t.copy_(saved)
        results.sort(key=lambda r: r[0])
        best = dict(results[0][1])
        self.apply(model, best)
        torch.cuda.synchronize(model.device)
        self.cache[key] = {"best": best, "ms": results[0][0],
                           "trials": [{"ms": ms, **cfg} for ms, cfg in results]}
        self.save()
        return best

    @staticmethod
    def apply(model: SimModelDO, cfg: dict) -> None:
        t = model.topo[0] if model.topo else None
        if t is not None and (t.constraint_order != "local" or t.chunk != cfg["chunk"]):
            model.set_constraint_order("local", cfg["chunk"])
        if cfg.get("parts") and getattr(model, "_parts_cfg", {}).get("tets_per_part") != cfg["parts"]:
            model.enable_parts(tets_per_part=cfg["parts"])
