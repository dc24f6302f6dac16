"""Serving loop: reset / step requests from many RL actors, continuously
batched onto one GPU.

Every admitted scene lives in a slot of a device arena (one ``SimModelDO``
whose environments are the slots; slots of one size class share the class
mesh's topology).  The loop

* takes requests as they arrive and keeps them in per-actor FIFOs;
* admits a new actor's scene when a slot of its class is free and the bytes
  of the resident scenes stay within the memory budget; otherwise the actor
  waits in the admission queue until a scene is closed;
* forms a tick from the oldest pending request of every resident actor that
  has no request on the GPU (at most ``max_batch``), and launches it without
  waiting for earlier ticks: ``ffsrl_apply_requests`` (resets, tool moves,
  grasp), ``ffsrl_frame_envs`` (one frame of the stepped scenes, refit
  included; ``ffsrl_frame_shared``, every CU on the tick's scenes, when the
  tick steps few scenes; ``ffsrl_frame_parts``, each scene in parts of
  ``parts`` tets with one tick in flight, when ``parts`` is given),
  ``ffsrl_render_envs`` (depth of every scene of the tick) and
  ``ffsrl_observe_envs`` (32 scalars).  Up to ``max_inflight`` ticks are on
  the GPU at once, on ``streams`` HIP streams, so requests that arrive while
  others run are admitted into the next tick and their work overlaps on the
  device;
* completes a tick when its HIP event has fired; a request's latency is the
  GPU completion time of its tick minus its arrival time.

Observations stay on the device: each slot owns its depth image and scalar
row, updated in place by the scene's latest request (the reference returns
zero-copy torch views of its state the same way, ``modelDO.py:1483-1518``).

Reset follows ``resetFixedModel`` (``modelDO.py:1589-1622``); a step moves the
instrument by the action's delta and optionally opens or closes the clamp
(``updateClampsDragKernel``, ``modelDO.py:247-279``).  The physics of a tick
is ``SimIntegratorDO.stepModel`` + ``SimBVH.refitBVH`` of the stepped scenes
(``integrator.py:1640-1744``, ``bvh.py:442-451``) and the render is
``renderNew`` in depth mode (``render.py:1099-1161``).
"""
from __future__ import annotations

import collections
import dataclasses
import gc
import time
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from ._lib import REQ_RESET, REQ_STEP, Requests, lib
from .model import SimModelDO
from .render import RaycastRendererDO
from .scenes import TetMesh

CLAMP_KEEP, CLAMP_OPEN, CLAMP_CLOSE = -1.0, 0.0, 1.0
CLOSE = 2  # request kind handled on the host: release the actor's scene
N_SCALARS = 32


@dataclasses.dataclass
class Request:
    actor: int
    kind: int                      # REQ_STEP, REQ_RESET or CLOSE
    action: tuple = (0.0, 0.0, 0.0, CLAMP_KEEP)
    t_arrival: float = 0.0         # seconds on the server clock
    seq: int = 0


@dataclasses.dataclass
class Completion:
    actor: int
    kind: int
    seq: int
    slot: int
    t_arrival: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


class _Tick:
    """Device and pinned-host buffers of one in-flight tick."""

    def __init__(self, max_batch: int, device):
        n = max_batch
        self.host = torch.zeros(8 * n, dtype=torch.float32, pin_memory=True)
        self.dev = torch.zeros(8 * n, dtype=torch.float32, device=device)
        self.event = torch.cuda.Event(enable_timing=True)
        self.items: List[tuple] = []
        self.busy = False
        self.n = 0


class SceneServer:
    """Continuous-batching server of scenes of a few size classes on one GPU.

    classes:        one mesh per size class (slots of a class share topology)
    capacity:       slots per class (the arena holds at most this many scenes)
    budget_bytes:   limit on the device bytes of the resident scenes
    model_kwargs:   SimModelDO physics keywords (substeps, constraint forms, ...)
    tips:           per-class initial instrument points, (n_class, 4, 3)
    """

    def __init__(self, classes: Sequence[TetMesh], capacity: Sequence[int], budget_bytes: float,
                 model_kwargs: dict, tips: np.ndarray, device="cuda:0", resolution: int = 64,
                 camera_pos=(0.0, 1.6, 5.5), look_at=(0.0, 0.45, 0.0), max_batch: int = 1024,
                 max_inflight: int = 2, streams: int = 2, block: int = 1024, grasp_radius: float = 0.05,
                 workspace: Optional[tuple] = None, shared: str = "auto", parts: int = 0):
        self.device = torch.device(device)
        self.classes = list(classes)
        self.capacity = [int(c) for c in capacity]
        slot_class = np.repeat(np.arange(len(self.classes)), self.capacity)
        tips = np.asarray(tips, np.float32).reshape(len(self.classes), 4, 3)
        self.model = SimModelDO(self.classes, slot_class.tolist(), device=device, tips=tips[slot_class],
                                **model_kwargs)
        m = self.model
        self.slot_class = slot_class
        self.n_slots = int(slot_class.size)
        self.renderer = RaycastRendererDO(device, m, cameraPos=camera_pos, resolution=resolution)
        self.renderer.lookAt(look_at)
        self.depth = self.renderer.depth                       # (n_slots, H, W), one image per slot
        self.obs = torch.zeros((self.n_slots, N_SCALARS), dtype=torch.float32, device=self.device)
        self.x0 = m.initialVertex
        self.tips0 = m.initialTips.reshape(-1).contiguous()
        self.base0 = m.initialBases.reshape(-1).contiguous()
        self.block = int(block)
        # shared-env frame for ticks that step at most half as many scenes as
        # the device has CUs ("auto"), always ("on") or never ("off")
        self.shared = shared
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        self.shared_max = {"auto": cus // 2, "on": self.n_slots, "off": 0}[shared]
        if shared == "auto" and lib.ffsrl_xcd_count() <= 0:   # no XCD map: one workgroup per scene
            self.shared_max = 0
        if self.shared_max > 0:
            m.enable_shared()
        # partitioned frame (tets per part; 0: off) for every tick: its parts
        # must be resident together, so the frames of two ticks must not run
        # at once: every tick goes to one stream (ticks in flight then queue
        # on the device while the host prepares the next)
        self.parts = int(parts)
        if self.parts:
            if int(streams) != 1:
                raise ValueError("the partitioned frame serves its ticks on one stream (streams=1)")
            m.enable_parts(tets_per_part=self.parts)
        self.radius = float(grasp_radius)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        if workspace is not None:
            self._lo = torch.tensor(workspace[0], dtype=torch.float32, device=self.device)
            self._hi = torch.tensor(workspace[1], dtype=torch.float32, device=self.device)
        else:
            self._lo = self._hi = None
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(streams)))]
        self.ticks = [_Tick(self.max_batch, self.device) for _ in range(self.max_inflight)]
        self.work = np.asarray(m.work, dtype=np.int64)         # LPT key of each slot
        self.class_bytes = [self._scene_bytes(k) for k in range(len(self.classes))]
        self.budget = float(budget_bytes)
        self.resident_bytes = 0
        self.free_slots: Dict[int, collections.deque] = {
            k: collections.deque(np.nonzero(slot_class == k)[0].tolist()) for k in range(len(self.classes))}
        self.actor_class: Dict[int, int] = {}
        self.actor_slot: Dict[int, int] = {}
        self.queues: Dict[int, collections.deque] = collections.defaultdict(collections.deque)
        self.waiting: collections.deque = collections.deque()   # actors awaiting admission
        self.inflight_actors = set()
        self.needs_reset = set()                                # admitted, state not yet reset
        self.ready: collections.deque = collections.deque()     # actors with pending requests, FIFO
        self.in_ready = set()
        self.completions: List[Completion] = []
        # on_complete(completion, server): server.depth[slot] and server.obs[slot]
        # hold the request's observation until that actor's next tick; work the
        # callback enqueues on torch's current stream (a copy of them) is
        # ordered before every later tick
        self.on_complete = None
        self.stats = collections.Counter()
        self._next_stream = 0
        self.t0 = None
        self.ev0 = None
        # the arena was filled on the current stream; the tick streams start after it
        torch.cuda.synchronize(self.device)

    # ------------------------------------------------------------ accounting
    def _scene_bytes(self, k: int) -> int:
        """Device bytes one scene of class k holds: x, p, v, x0 (float4),
        drag, BVH bounds, tool points, depth image and observation row."""
        m = self.classes[k]
        nodes = self.model.topo[k].num_nodes
        return (m.num_vertices * (4 * 16 + 4) + nodes * 32 + 12 * 4 * 2
                + self.depth.shape[1] * self.depth.shape[2] * 4 + N_SCALARS * 4)

    def register(self, actor: int, cls: int) -> None:
        self.actor_class[int(actor)] = int(cls)

    def _admit(self, actor: int) -> bool:
        k = self.actor_class[actor]
        if not self.free_slots[k] or self.resident_bytes + self.class_bytes[k] > self.budget:
            return False
        self.actor_slot[actor] = self.free_slots[k].popleft()
        self.resident_bytes += self.class_bytes[k]
        self.needs_reset.add(actor)
        self.stats["admitted"] += 1
        return True

    def _release(self, actor: int) -> None:
        slot = self.actor_slot.pop(actor)
        k = self.actor_class[actor]
        self.free_slots[k].append(slot)
        self.resident_bytes -= self.class_bytes[k]
        self.stats["closed"] += 1
        while self.waiting and self._admit(self.waiting[0]):
            a = self.waiting.popleft()
            self._mark_ready(a)

    # ------------------------------------------------------------- requests
    def start_clock(self) -> None:
        torch.cuda.synchronize(self.device)
        self.ev0 = torch.cuda.Event(enable_timing=True)
        self.ev0.record(torch.cuda.current_stream(self.device))
        torch.cuda.synchronize(self.device)
        self.t0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.t0

    def _mark_ready(self, actor: int) -> None:
        if actor not in self.in_ready and self.queues[actor]:
            self.ready.append(actor)
            self.in_ready.add(actor)

    def submit(self, req: Request) -> None:
        a = int(req.actor)
        if a not in self.actor_class:
            raise KeyError(f"actor {a} has no registered scene class")
        self.queues[a].append(req)
        if a not in self.actor_slot:
            if a not in self.waiting and not self._admit(a):
                self.waiting.append(a)
                self.stats["admission_waits"] += 1
                return
            if a in self.waiting:
                return
        self._mark_ready(a)

    # ----------------------------------------------------------------- ticks
    def _free_tick(self) -> Optional[_Tick]:
        for t in self.ticks:
            if not t.busy:
                return t
        return None

    def launch(self) -> int:
        """Launch one tick from the ready actors; returns its request count."""
        tick = self._free_tick()
        if tick is None or not self.ready:
            return 0
        items = []
        deferred = []
        while self.ready and len(items) < self.max_batch:
            a = self.ready.popleft()
            self.in_ready.discard(a)
            if a in self.inflight_actors:
                deferred.append(a)
                continue
            req = self.queues[a][0]
            if req.kind == CLOSE:
                self.queues[a].popleft()
                self._release(a)
                self._complete_now(req)
                self._mark_ready(a)
                continue
            self.queues[a].popleft()
            if a in self.needs_reset:
                # a slot holds the previous occupant's state until its first reset
                if req.kind == REQ_STEP:
                    req = dataclasses.replace(req, kind=REQ_RESET)
                    self.stats["step_as_reset"] += 1
                self.needs_reset.discard(a)
            items.append((a, self.actor_slot[a], req))
        for a in deferred:
            self._mark_ready(a)
        if not items:
            return 0
        n = len(items)
        envs = np.fromiter((s for _, s, _ in items), np.int32, n)
        kinds = np.fromiter((r.kind for _, _, r in items), np.int32, n)
        acts = np.asarray([r.action for _, _, r in items], np.float32).reshape(n, 4)
        step = envs[kinds == REQ_STEP]
        step = step[np.argsort(-self.work[step], kind="stable")].astype(np.int32)   # largest first
        m = int(step.size)
        h = tick.host
        hi = h.view(torch.int32).numpy()
        hf = h.numpy()
        hi[0:n] = envs
        hi[n:2 * n] = kinds
        hi[2 * n:2 * n + m] = step
        off = 4 * self.max_batch
        hf[off:off + 4 * n] = acts.reshape(-1)
        stream = self.streams[self._next_stream % len(self.streams)]
        self._next_stream += 1
        model = self.model
        # completion callbacks copy a slot's outputs on the current stream: the
        # tick that overwrites them runs after those copies
        if self.on_complete is not None:
            stream.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(stream):
            tick.dev.copy_(tick.host, non_blocking=True)
            dv = tick.dev.view(torch.int32)
            p_env, p_kind, p_step = dv.data_ptr(), dv.data_ptr() + 4 * n, dv.data_ptr() + 8 * n
            p_act = tick.dev.data_ptr() + 4 * off
            s = stream.cuda_stream
            r = Requests()
            r.n, r.envs, r.kind, r.action = n, p_env, p_kind, p_act
            r.x0, r.tips0, r.base0 = self.x0.data_ptr(), self.tips0.data_ptr(), self.base0.data_ptr()
            r.lo = None if self._lo is None else self._lo.data_ptr()
            r.hi = None if self._hi is None else self._hi.data_ptr()
            lib.ffsrl_apply_requests(model.batch, r, self.radius, s)
            if m and self.parts:
                lib.ffsrl_frame_parts(model.batch, model.params, model.parts, p_step, m, 1, s)
            elif m and m <= self.shared_max:
                lib.ffsrl_frame_shared(model.batch, model.params, p_step, m, self.block, 1, s)
            elif m:
                lib.ffsrl_frame_envs(model.batch, model.params, p_step, m, self.block, 1, s)
            cam = self.renderer.camera
            lib.ffsrl_render_envs(model.batch, cam, p_env, n, self.depth.data_ptr(), None, None, None, None, None,
                                  None, s)
            lib.ffsrl_observe_envs(model.batch, p_env, n, self.x0.data_ptr(), float(model.groundLevel),
                                   self.obs.data_ptr(), s)
            tick.event.record(stream)
        tick.items = items
        tick.busy = True
        tick.n = n
        for a, _, _ in items:
            self.inflight_actors.add(a)
        self.stats["ticks"] += 1
        self.stats["tick_requests"] += n
        return n

    def _complete_now(self, req: Request) -> None:
        c = Completion(req.actor, req.kind, req.seq, -1, req.t_arrival, self.now() if self.t0 else 0.0)
        self.completions.append(c)
        if self.on_complete:
            self.on_complete(c, None)

    def poll(self, block: bool = False) -> int:
        """Complete every tick whose event has fired (or all, if block)."""
        done = 0
        for t in self.ticks:
            if not t.busy:
                continue
            if block:
                t.event.synchronize()
            elif not t.event.query():
                continue
            t_done = self.ev0.elapsed_time(t.event) * 1e-3 if self.ev0 is not None else 0.0
            for a, slot, req in t.items:
                self.inflight_actors.discard(a)
                c = Completion(a, req.kind, req.seq, slot, req.t_arrival, t_done)
                self.completions.append(c)
                self.stats["step" if req.kind == REQ_STEP else "reset"] += 1
                if self.on_complete:
                    self.on_complete(c, self)
                self._mark_ready(a)
            t.items = []
            t.busy = False
            done += 1
        return done

    def run_until_idle(self) -> None:
        """Launch and complete ticks until no request is pending (tests and
        offline drivers; the timed loop is ``serve``)."""
        while True:
            launched = self.launch()
            if not launched:
                if not any(t.busy for t in self.ticks):
                    if not self.ready:
                        break
                self.poll(block=True)
            else:
                self.poll()

    def idle(self) -> bool:
        return not self.ready and not any(t.busy for t in self.ticks)


# ---------------------------------------------------------------- load driver
@dataclasses.dataclass
class ServingLoad:
    """Seeded Poisson stream of step / reset requests from logical actors.

    Each actor's scene class is drawn once with probabilities ``class_p``; an
    actor's first request is a reset, later ones are resets with probability
    ``reset_p`` and steps otherwise.  A step's action is a Gaussian tool move
    (``sigma`` per axis) and closes the clamp on the first step after a reset.
    """
    rate: float = 20000.0
    actors: int = 4096
    class_p: tuple = (0.4, 0.3, 0.2, 0.1)
    reset_p: float = 0.01
    sigma: float = 0.002
    seed: int = 0

    def generate(self, duration: float):
        rng = np.random.default_rng(self.seed)
        cls = rng.choice(len(self.class_p), size=self.actors, p=np.asarray(self.class_p) / sum(self.class_p))
        n = int(rng.poisson(self.rate * duration))
        t = np.sort(rng.uniform(0.0, duration, size=n))
        who = rng.integers(0, self.actors, size=n)
        u = rng.uniform(size=n)
        moves = rng.normal(0.0, self.sigma, size=(n, 3)).astype(np.float32)
        started = np.zeros(self.actors, bool)
        fresh = np.zeros(self.actors, bool)
        reqs = []
        for i in range(n):
            a = int(who[i])
            if not started[a] or u[i] < self.reset_p:
                started[a] = True
                fresh[a] = True
                reqs.append(Request(a, REQ_RESET, (0.0, 0.0, 0.0, CLAMP_KEEP), float(t[i]), i))
            else:
                clamp = CLAMP_CLOSE if fresh[a] else CLAMP_KEEP
                fresh[a] = False
                reqs.append(Request(a, REQ_STEP, (float(moves[i, 0]), float(moves[i, 1]), float(moves[i, 2]), clamp),
                                    float(t[i]), i))
        return cls, reqs


def serve(server: SceneServer, requests: Sequence[Request], duration: float, min_batch: int = 1,
          window: float = 0.0) -> dict:
    """Replays ``requests`` in real time (arrival times on the server clock)
    for ``duration`` seconds and returns the latency / throughput summary.
    A tick is launched when a tick buffer is free and either ``min_batch``
    requests are ready or the oldest ready request has waited ``window`` s."""
    gc_was = gc.isenabled()
    gc.collect()
    gc.disable()                 # no collector pauses inside the timed loop
    server.start_clock()
    i, n = 0, len(requests)
    while True:
        now = server.now()
        while i < n and requests[i].t_arrival <= now:
            server.submit(requests[i])
            i += 1
        server.poll()
        if server.ready:
            head = server.queues[server.ready[0]]
            waited = now - head[0].t_arrival if head else window
            if len(server.ready) >= min_batch or waited >= window:
                server.launch()
        if i >= n or now >= duration + 5.0:        # stream replayed (or drain limit)
            break
    while not server.idle():
        server.launch()
        server.poll(block=True)
    wall = server.now()
    if gc_was:
        gc.enable()
    out = summarize(server.completions, wall, duration)
    st = server.model.frame_status().cpu().numpy()
    out["status_scenes"] = int((st != 0).sum())      # slots whose frames reported a status bit
    return out


def summarize(completions: Sequence[Completion], wall: float, duration: float) -> dict:
    lat = np.asarray([c.latency for c in completions if c.kind in (REQ_STEP,)], np.float64)
    lat_all = np.asarray([c.latency for c in completions if c.kind in (REQ_STEP, REQ_RESET)], np.float64)
    steps = int(lat.size)
    out = dict(completed_steps=steps, completed_resets=int(lat_all.size - steps), wall_s=wall,
               env_steps_per_s=steps / wall if wall > 0 else 0.0)
    if lat.size:
        out.update(p50_step_ms=float(np.percentile(lat, 50) * 1e3), p99_step_ms=float(np.percentile(lat, 99) * 1e3),
                   max_step_ms=float(lat.max() * 1e3))
    if lat_all.size:
        out.update(p50_ms=float(np.percentile(lat_all, 50) * 1e3), p99_ms=float(np.percentile(lat_all, 99) * 1e3))
    return out
