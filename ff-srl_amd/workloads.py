"""Synthetic workloads of the benchmark configurations.

* cube     - 1 env, 10x10x10 vertex cube, 5 tets/cell (3645 tets), XPBD
             distance (compliance 1e-6) + XPBD volume (compliance 0), bottom
             face pinned, g = -9.81, dt = 1/60, 10 substeps x 1 iteration,
             seeded initial velocity noise U[-0.01, 0.01], 128x128 depth.
* ellipsoids - n envs of a seeded ~20k-tet ellipsoid (1, 0.6, 0.4), Neo-Hookean
             (mu 1e4, lambda 5e4), bottom 10% pinned, 8 grasped surface
             vertices pulled toward a tool tip on a random walk (sigma 2 mm),
             10 substeps, 64x64 depth + RGB.
* mixed    - n envs (1024) with tets log-uniform in [1k, 40k], ellipsoid or
             box (50/50), distance + volume + Neo-Hookean, grasp of the
             surface vertex nearest the tool tip in 50% of envs, 10
             substeps, refit + 64x64 depth.
* large    - n envs (8192) of the 20k-tet ellipsoid, Neo-Hookean + volume +
             grasp of one surface vertex, 20 substeps x 2 iterations,
             128x128 RGB-D.
Every mesh is generated from its own seed; environments of the mixed workload
each get a distinct mesh.

Grasp size.  The grasp (dragConstraintsDO, integrator.py:412-442) pulls every
grasped vertex toward one point, the tool tip.  Eight grasped vertices
collapse onto it within a frame; with the stiff XPBD distance (compliance
1e-6) and volume (compliance 0) constraints of the mixed workload (volume
and Neo-Hookean in the large one), the collapsed tets then drive the scheme
apart: the float32 and float64 oracles differ by up to 2 mesh diagonals
after one frame of the mixed workload.  One grasped vertex keeps both
workloads reproducible (the two oracles agree to 7e-6 of the diagonal over
3 frames of the mixed workload and 1 frame of the large one), so those
workloads grasp one vertex; the Neo-Hookean-only ellipsoid workload keeps
its eight (`grasp_k`).
"""
from __future__ import annotations

import dataclasses
import multiprocessing as mp
import os
from typing import List, Optional

import numpy as np

from . import scenes as sc

NH_PARAMS = dict(neoHookean=True, globalVolumeCompliance=1.0 / 5e4, globalDeviatoricCompliance=1.0 / 1e4)
XPBD_DV = dict(distance="xpbd", volume="xpbd", distanceCompliance=1e-6, volumeComplianceXPBD=0.0)
BASE = dict(simFrameRate=60, environmentGravity=(0.0, -9.81, 0.0), environmentVelocityDampening=0.1,
            environmentGroundLevel=0.0, globalKsDrag=1.0)


@dataclasses.dataclass
class Workload:
    name: str
    meshes: list
    env_mesh: list
    model_kwargs: dict
    resolution: int
    camera_pos: tuple
    look_at: tuple
    tips: Optional[np.ndarray] = None
    grasp: Optional[List[Optional[np.ndarray]]] = None
    initial_velocity: Optional[np.ndarray] = None
    walk_sigma: float = 0.002
    rgb: bool = False
    caps_base: Optional[np.ndarray] = None     # (n_env, 4, 3) instrument capsule bases
    caps_radius: Optional[np.ndarray] = None   # (n_env, 4) capsule radii
    laparoscopes: Optional[list] = None        # instrument kinematics per env (replaces tips / capsules)
    data: str = "synthetic (seeded meshes, seeded tool random walk)"
    setup: Optional[object] = None             # callable(model) run once after the model is built

    @property
    def num_envs(self):
        return len(self.env_mesh)

    def totals(self):
        T = sum(self.meshes[k].num_tets for k in self.env_mesh)
        E = sum(self.meshes[k].num_edges for k in self.env_mesh)
        V = sum(self.meshes[k].num_vertices for k in self.env_mesh)
        F = sum(self.meshes[k].num_tris for k in self.env_mesh)
        return dict(tets=T, edges=E, vertices=V, tris=F)


def _gen(args):
    shape, tets, seed = args
    return sc.solid(shape, tets, seed=seed)


def generate(specs, workers: Optional[int] = None):
    """Generate meshes for (shape, tets, seed) specs on a process pool: forked
    before the GPU runtime is initialised, spawned after it (a forked child
    of a process with a live GPU runtime is unsafe).  FFSRL_GEN_WORKERS sets
    the pool size (1: generate in this process, e.g. under a profiler whose
    preloaded library should not run in children).  The pool is closed and
    joined, so its workers exit on their own rather than by a signal."""
    workers = workers or int(os.environ.get("FFSRL_GEN_WORKERS", "0")) or min(16, os.cpu_count() or 1)
    if workers <= 1 or len(specs) <= 8:
        return [_gen(s) for s in specs]
    live = False
    try:
        import torch
        live = torch.cuda.is_initialized()
    except Exception:
        pass
    pool = mp.get_context("spawn" if live else "fork").Pool(min(workers, len(specs)))
    try:
        out = pool.map(_gen, specs, chunksize=4)
        pool.close()
        pool.join()
        return out
    except BaseException:
        pool.terminate()
        raise


def _tool(mesh, rng_offset=0.02, k=8):
    tip = sc.top_point(mesh) + np.array([0.0, rng_offset, 0.0], np.float32)
    return np.repeat(tip[None], 4, axis=0).astype(np.float32), sc.nearest_surface_vertices(mesh, tip, k)


# Synthetic instrument around the tool point t (the four instrument points
# all sit at t): a rod of radius 3 cm reaching 0.5 up from t and two jaws of
# radius 1.5 cm closing on t from 3 cm to either side; point 3 (the drag
# point) is a capsule of radius 0, as in laparoscope.py:375-381.
INSTRUMENT_BASE_OFFSET = np.array([[0.0, 0.5, 0.0], [-0.03, 0.08, 0.0], [0.03, 0.08, 0.0], [0.0, 0.0, 0.0]],
                                  np.float32)
INSTRUMENT_RADII = np.array([0.03, 0.015, 0.015, 0.0], np.float32)


def _capsules(tips: np.ndarray):
    tips = np.asarray(tips, np.float32).reshape(-1, 4, 3)
    return tips + INSTRUMENT_BASE_OFFSET[None], np.tile(INSTRUMENT_RADII, (tips.shape[0], 1))


def cube(seed: int = 0) -> Workload:
    mesh = sc.cube_grid(10)
    rng = np.random.default_rng(seed)
    v0 = rng.uniform(-0.01, 0.01, size=(mesh.num_vertices, 3)).astype(np.float32)
    kw = dict(BASE, simSubsteps=10, simConstraintsSteps=1, **XPBD_DV)
    return Workload("cube-1env-3645tets", [mesh], [0], kw, 128, (2.2, 1.8, 2.6), (0.5, 0.45, 0.5),
                    tips=np.zeros((1, 4, 3), np.float32), initial_velocity=v0)


def ellipsoids(n_env: int = 64, tets: int = 20000, seed: int = 0, substeps=10, iterations=1, resolution=64,
               volume: bool = False, shared: bool = False, grasp_k: int = 8) -> Workload:
    specs = [("ellipsoid", tets, seed + (0 if shared else i)) for i in range(1 if shared else n_env)]
    meshes = generate(specs)
    tips, grasp = [], []
    for i in range(n_env):
        t, g = _tool(meshes[0 if shared else i], k=grasp_k)
        tips.append(t)
        grasp.append(g)
    kw = dict(BASE, simSubsteps=substeps, simConstraintsSteps=iterations, distance=None, volume=None, **NH_PARAMS)
    if volume:
        kw.update(volume="xpbd", volumeComplianceXPBD=0.0)
    env_mesh = [0] * n_env if shared else list(range(n_env))
    cb, cr = _capsules(np.stack(tips))
    return Workload(f"ellipsoid-{n_env}env-{meshes[0].num_tets}tets", meshes, env_mesh, kw, resolution,
                    (0.0, 1.6, 5.5), (0.0, 0.45, 0.0), tips=np.stack(tips), grasp=grasp, rgb=True,
                    caps_base=cb, caps_radius=cr)


def mixed(n_env: int = 1024, lo: int = 1000, hi: int = 40000, seed: int = 0, substeps=10, iterations=1,
          resolution=64, grasp_k: int = 1) -> Workload:
    sizes, grasp_on = sc.mixed_sizes(n_env, lo, hi, seed)
    specs = [(shape, t, seed * 100003 + i) for i, (shape, t) in enumerate(sizes)]
    meshes = generate(specs)
    tips, grasp = [], []
    for i, m in enumerate(meshes):
        t, g = _tool(m, k=grasp_k)
        tips.append(t)
        grasp.append(g if grasp_on[i] else None)
    kw = dict(BASE, simSubsteps=substeps, simConstraintsSteps=iterations, **XPBD_DV, **NH_PARAMS)
    cb, cr = _capsules(np.stack(tips))
    return Workload(f"mixed-{n_env}env-logU[{lo},{hi}]tets", meshes, list(range(n_env)), kw, resolution,
                    (0.0, 1.6, 5.5), (0.0, 0.45, 0.0), tips=np.stack(tips), grasp=grasp, caps_base=cb, caps_radius=cr)


# SimModelDO arguments and set-up of the reference's liver retraction test
# (tests/testLiverRetraction.py:29-58); its DO step runs the PBD forms.
LIVER_KW = dict(simFrameRate=30, simSubsteps=8, simConstraintsSteps=1, globalKsDrag=1.0,
                globalLaparoscopeDragLookupRadius=1.0, globalKsDistance=1.0, globalKsVolume=1.0, distance="pbd",
                volume="pbd", environmentGravity=(0.0, -9.8, 0.0), workspaceLow=(-7.0, 0.7, -10.6),
                workspaceHigh=(7.0, 9.0, 7.4))
LIVER_RCM, LIVER_RCM_ROT, LIVER_EFFECTOR = (9.036183, 35.260103, 1.567807), (0.0, 0.0, 1.0), (0.0, 10.0, 4.0)


def liver(path: str, n_env: int = 1024, resolution: int = 64) -> Workload:
    """n_env copies of a USD scene (the reference's liver retraction scene, or
    its converted .npz), with the instrument placed as in the reference's
    test: rotation centre, then the drag point moved to (0, 10, 4).  The tool
    random walk drives the instrument through Cartesian actions (0.05 per
    frame)."""
    from . import usd_scene
    sc = usd_scene.scene(path)
    mesh = usd_scene.merged_mesh(sc)

    def setup(model):
        import torch
        n = model.numEnvs
        dev = model.device
        model.setRotationCenterInitialPosition(torch.tensor([LIVER_RCM] * n, device=dev),
                                               torch.tensor([LIVER_RCM_ROT] * n, device=dev))
        model.setEffectorInitialPosition(torch.tensor(LIVER_EFFECTOR, device=dev) - model.getLaparoscopePositionsTensor())

    name = f"liver-retraction-{n_env}env-{mesh.num_tets}tets"
    return Workload(name, [mesh], [0] * n_env, dict(LIVER_KW), resolution, (0.0, 50.0, 50.0), (0.0, 10.0, 0.0),
                    laparoscopes=[sc.laparoscope] * n_env, walk_sigma=0.05, setup=setup,
                    data=f"the reference's scene ({os.path.basename(path)}), seeded Cartesian tool random walk")


def large(n_env: int = 8192, tets: int = 20000, seed: int = 0, grasp_k: int = 1) -> Workload:
    w = ellipsoids(n_env, tets, seed, substeps=20, iterations=2, resolution=128, volume=True, shared=True,
                   grasp_k=grasp_k)
    w.name = f"ellipsoid-{n_env}env-{w.meshes[0].num_tets}tets-S20I2"
    return w


@dataclasses.dataclass
class ServingWorkload:
    name: str
    classes: list                  # one mesh per scene size class
    class_p: tuple
    tips: np.ndarray               # (n_class, 4, 3) initial instrument points
    model_kwargs: dict
    resolution: int
    camera_pos: tuple
    look_at: tuple
    rate: float
    actors: int


def serving(actors: int = 4096, rate: float = 20000.0, tets=(2000, 8000, 20000, 40000),
            class_p=(0.4, 0.3, 0.2, 0.1), seed: int = 0, substeps=10, resolution=64) -> ServingWorkload:
    """Scenes of four sizes (ellipsoids, seeded), distance + volume +
    Neo-Hookean, 10 substeps, 64x64 depth; Poisson requests from `actors`."""
    meshes = generate([("ellipsoid", t, seed * 7919 + i) for i, t in enumerate(tets)])
    tips = np.stack([_tool(m)[0] for m in meshes])
    kw = dict(BASE, simSubsteps=substeps, simConstraintsSteps=1, **XPBD_DV, **NH_PARAMS)
    return ServingWorkload(f"serving-{actors}actors-{int(rate)}rps-tets{'/'.join(str(t) for t in tets)}", meshes,
                           tuple(class_p), tips, kw, resolution, (0.0, 1.6, 5.5), (0.0, 0.45, 0.0), rate, actors)


def build_model(pkg, w: Workload, device="cuda:0"):
    """SimModelDO for a workload, with the grasp weights set."""
    model = pkg.SimModelDO(w.meshes, w.env_mesh, device=device, tips=w.tips, initialVelocity=w.initial_velocity,
                           capsuleBases=w.caps_base, capsuleRadii=w.caps_radius, laparoscopes=w.laparoscopes,
                           **w.model_kwargs)
    if w.setup is not None:
        w.setup(model)
    if w.grasp is not None:
        import torch
        ids = [np.asarray(g, np.int64) + int(model.env_vbase[i]) for i, g in enumerate(w.grasp) if g is not None]
        if ids:
            model.drag[torch.as_tensor(np.concatenate(ids), device=model.device)] = 1.0
    return model
