"""Benchmark: env-steps/s at 1024 mixed-size environments, p50/p99 step
latency, batched vs plain, on one MI355X.

An env-step is one frame of one environment: S substeps x I iterations of the
constraint projection, the BVH refit and a depth render.  A bench step is one
frame of all environments: the tool move, the batched frame kernel (physics +
refit) and the depth render.  Inputs are resident in HBM before timing.

Prints one JSON line (rank 0, one GPU).
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic():
    """HBM bytes per frame-kernel launch measured with rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE (calibrated on a device copy), by workload:
    each summary under profiles/ states its figures on lines
    `traffic-per-launch: <workload> <kernel> <bytes> bytes (measured <date>)`.
    Returns {(workload, kernel): (bytes, source)}."""
    out = {}
    pat = re.compile(r"^traffic-per-launch:\s+(\S+)\s+(\S+)\s+([0-9.eE+]+) bytes \(measured ([^)]*)\)")
    for name in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(PROFILES, name)):
            m = pat.match(line.strip().strip("`"))
            if m:
                out[(m.group(1), m.group(2))] = (float(m.group(3)), f"profiles/{name}, measured {m.group(4)}")
    return out


def one_launch_count(model, n):
    """Kernel launches of one ffsrl_frame_do frame (with refit) over n envs."""
    per_it = ((model.vol_mode != 0) + (model.dist_mode != 0) + bool(model.nh_on) + 2 +
              (2 if model.collision else 0))
    reset = model.carries_lambda and (not model.lambda_legacy or model.simConstraints > 1)
    lam = (model.lam_e is not None) + (model.lam_t is not None) if reset else 0
    groups = max(int(t.num_groups) for t in model.topo)
    return model.simSubsteps * (2 + lam + model.simConstraints * per_it) + groups


def copy_peak_gbs(dev, n_bytes=1 << 30, reps=10):
    """Device-to-device copy bandwidth on this GPU (read + write bytes / time,
    best of `reps`), the practical ceiling beside the nominal 8 TB/s."""
    import torch
    a = torch.ones(n_bytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del a, b
    torch.cuda.empty_cache()
    return 2.0 * n_bytes / (best * 1e-3) / 1e9


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="mixed", choices=["mixed", "ellipsoids", "cube", "large", "serve", "liver"])
    ap.add_argument("--scene", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                                     "liver_retraction.npz"),
                    help="USD crate scene or converted .npz for --workload liver")
    ap.add_argument("--resolution", type=int, default=None, help="render resolution override")
    ap.add_argument("--envs", type=int, default=None)
    ap.add_argument("--block", type=int, default=1024)
    ap.add_argument("--shared", choices=["auto", "on", "off"], default="auto",
                    help="shared-env frame: autotuned (auto), forced on, or off")
    ap.add_argument("--fold", choices=["on", "off"], default="on",
                    help="the shared frame's folded form for envs of at most 5,104 vertices (same results)")
    ap.add_argument("--parts", choices=["on", "off"], default="off",
                    help="partitioned frame: each env cut into parts, one workgroup per part (same results)")
    ap.add_argument("--tets-per-part", type=int, default=5120,
                    help="constraints per part of the partitioned frame (1073741824: as few parts as LDS allows)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="colour-run length of the constraint order with --no-autotune (0: the model's default)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --block and largest-first dispatch instead of the autotuned launch shape")
    ap.add_argument("--plain-frames", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the multi-core CPU baseline")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--gen-workers", type=int, default=0,
                    help="processes generating the meshes (1: in this process, as under a profiler; 0: up to 16)")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="length of the serving run after the batched measurement (0: skip); the serving "
                         "configuration is 60 s")
    ap.add_argument("--serve-rate", type=float, default=20000.0)
    ap.add_argument("--serve-actors", type=int, default=4096)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-streams", type=int, default=1)
    ap.add_argument("--serve-inflight", type=int, default=1)
    ap.add_argument("--serve-parts", type=int, default=1280,
                    help="serve every tick with the partitioned frame, this many tets per part (one tick in flight); "
                         "0: the shared-env frame (--serve-inflight 2 --serve-streams 2 measured best with it)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device-to-device copy before timing (HBM counter calibration)")
    ap.add_argument("--collision", action="store_true",
                    help="add the instrument collision pass to every constraint iteration (not the headline)")
    return ap.parse_args()


def make_workload(wl, args):
    if args.workload == "mixed":
        w = wl.mixed(args.envs or 1024, seed=args.seed)
    elif args.workload == "ellipsoids":
        w = wl.ellipsoids(args.envs or 64, seed=args.seed)
    elif args.workload == "large":
        w = wl.large(args.envs or 8192, seed=args.seed)
    elif args.workload == "liver":
        w = wl.liver(args.scene, args.envs or 1024)
    else:
        w = wl.cube(args.seed)
    if args.resolution:
        w.resolution = args.resolution
    if args.collision:
        if w.caps_base is None and w.laparoscopes is None:
            raise SystemExit(f"workload {w.name} has no instrument capsules")
        w.model_kwargs = dict(w.model_kwargs, collision=True, globalKsContact=1.0)
        w.name += "+collision"
    return w


def algorithmic_bytes(w, model):
    """Compulsory bytes of one frame kernel launch (all envs), per env:
    S x [predict 40 V + I x (28 V + 12 E + T x (56 with Neo-Hookean, else 20))
    + integrate 48 V] + 8 B per lambda-carrying constraint per extra iteration
    + refit (12 V_render + 16 F + 40 N); with the instrument collision,
    I x (24 F + 32 V) more per substep (triangle ids, surface positions, and
    one more read and write of every position)."""
    kw = w.model_kwargs
    S, I = kw["simSubsteps"], kw["simConstraintsSteps"]
    nh = kw.get("neoHookean", False)
    has_vol = kw.get("volume") is not None
    has_dist = kw.get("distance") is not None
    total = 0
    for e, k in enumerate(w.env_mesh):
        m = w.meshes[k]
        topo = model.topo[k]
        V, E, T, F, N = m.num_vertices, m.num_edges, m.num_tets, m.num_tris, topo.num_nodes
        t_bytes = 56 if nh else (20 if has_vol else 0)
        proj = 28 * V + (12 * E if has_dist else 0) + t_bytes * T
        lam = 0
        if I > 1:
            lam = 8 * ((E if kw.get("distance") == "xpbd" else 0) + (T if kw.get("volume") == "xpbd" else 0))
        if kw.get("collision"):
            proj += 24 * F + 32 * V
        per_sub = 40 * V + I * proj + (I - 1) * lam + 48 * V
        vr = m.surface_vertices().size
        total += S * per_sub + 12 * vr + 16 * F + 40 * N
    return total


def cpu_baseline(pkg, orc, w, model, ren, seconds, seed, threads):
    """The oracle (fp32 serial C restatement) on a random sample of the same
    envs: one frame (physics + refit + depth, + normals and RGB when the
    workload renders colour) per sampled env.  Envs are independent, so the
    multi-core figure runs one env per host thread (the library releases the
    interpreter lock); the 1-core figure runs them one after another."""
    import concurrent.futures
    import threading
    rng = np.random.default_rng(seed + 17)
    order = list(rng.permutation(w.num_envs))
    kw = w.model_kwargs
    cfg = orc.StepConfig(substeps=kw["simSubsteps"], iterations=kw["simConstraintsSteps"],
                         fps=kw["simFrameRate"], gravity=kw.get("environmentGravity", (0.0, -9.8, 0.0)),
                         damping=kw.get("environmentVelocityDampening", 0.1),
                         dist_mode={None: 0, "pbd": 1, "xpbd": 2}[kw.get("distance")],
                         vol_mode={None: 0, "pbd": 1, "xpbd": 2}[kw.get("volume")],
                         nh_on=int(kw.get("neoHookean", False)),
                         nh_vol_c=kw.get("globalVolumeCompliance", 1.0),
                         nh_dev_c=kw.get("globalDeviatoricCompliance", 1.0),
                         dist_k=kw.get("globalKsDistance", 1.0), vol_k=kw.get("globalKsVolume", 1.0),
                         dist_compliance=kw.get("distanceCompliance", 0.0) if kw.get("distance") == "xpbd" else None,
                         vol_compliance=kw.get("volumeComplianceXPBD", 0.0) if kw.get("volume") == "xpbd" else None,
                         drag_ks=kw.get("globalKsDrag", 1.0),
                         collide=int(kw.get("collision", False)), contact_k=kw.get("globalKsContact", 1.0))
    dirs = ren.dirs.cpu().numpy()
    origin = np.asarray(ren.camera.origin[:], np.float32)
    light = np.asarray(ren.camera.light_dir[:], np.float32)
    bvhs, tex = {}, {}
    for k in set(w.env_mesh[e] for e in order[:4096]):
        t = model.topo[k]
        bvhs[k] = {"left": t.left, "first": t.first, "count": t.count, "level": t.level, "tri_map": t.tri_map}
        if w.rgb:
            tex[k] = model.texture_inputs(k)

    def one(e):
        k = w.env_mesh[e]
        m = w.meshes[k]
        sim = orc.OracleSim([m], [0], np.float32, tips=w.tips[e:e + 1] if w.tips is not None else None,
                            initial_velocity=w.initial_velocity,
                            caps_base=w.caps_base[e:e + 1] if w.caps_base is not None else None,
                            caps_radius=w.caps_radius[e:e + 1] if w.caps_radius is not None else None)
        if w.grasp is not None and w.grasp[e] is not None:
            sim.drag[w.grasp[e]] = 1.0
        sim.step(cfg, 1)
        bvh = bvhs[k]
        mn, mx = orc.refit(bvh, m.tris, sim.x)
        orc.render_depth(ren.height, ren.width, origin, dirs, bvh, mn, mx, m.tris, sim.x)
        if w.rgb:
            nrm = orc.vertex_normals(m.tris, sim.x)
            orc.render_rgb(ren.height, ren.width, 1, origin, dirs, bvh, mn, mx, m.tris, sim.x, nrm,
                           vcolor=m.colors, light=light, intensity=ren.camera.light_intensity, **tex[k])
        return m.num_tets

    def run(workers, budget, envs):
        it = iter(envs)
        lock = threading.Lock()
        done, tets = [0], [0]
        t0 = time.perf_counter()

        def worker():
            while time.perf_counter() - t0 < budget:
                with lock:
                    e = next(it, None)
                if e is None:
                    return
                nt = one(e)
                with lock:
                    done[0] += 1
                    tets[0] += nt

        with concurrent.futures.ThreadPoolExecutor(workers) as ex:
            for f in [ex.submit(worker) for _ in range(workers)]:
                f.result()
        el = time.perf_counter() - t0
        mean_t_sample = tets[0] / max(done[0], 1)
        mean_t_pop = sum(w.meshes[k].num_tets for k in w.env_mesh) / w.num_envs
        return done[0] / el * (mean_t_sample / mean_t_pop), done[0], el, mean_t_sample, mean_t_pop

    what = f"physics + refit + {ren.height}x{ren.width} depth{' + normals + RGB' if w.rgb else ''}"
    r1, d1, el1, ms1, mp = run(1, seconds / 3.0, order)
    rN, dN, elN, msN, _ = run(threads, 2.0 * seconds / 3.0, order[d1:] + order[:d1])
    one_core = {"value": r1, "unit": "env-steps/s", "cores": 1, "kind": "port",
                "sample": f"oracle fp32 serial C restatement, {d1} randomly sampled envs x 1 frame ({what}) "
                          f"in {el1:.1f} s, scaled by mean tets {ms1:.0f}/{mp:.0f}"}
    all_cores = {"value": rN, "unit": "env-steps/s", "cores": threads, "kind": "port",
                 "sample": f"oracle fp32 C restatement, one env per host thread, {dN} randomly sampled envs x "
                           f"1 frame ({what}) in {elN:.1f} s on {threads} threads, scaled by mean tets "
                           f"{msN:.0f}/{mp:.0f}"}
    return all_cores, one_core


def serving_run(pkg, wl, args, dev):
    """Poisson stream of step / reset requests from RL actors, continuously
    batched (ffsrl_apply_requests + ffsrl_frame_envs + ffsrl_render_envs +
    ffsrl_observe_envs per tick); latency is arrival -> GPU completion."""
    import torch
    S = pkg.serve
    sw = wl.serving(actors=args.serve_actors, rate=args.serve_rate, seed=args.seed)
    load = S.ServingLoad(rate=sw.rate, actors=sw.actors, class_p=sw.class_p, seed=args.seed)
    cls, reqs = load.generate(args.serve_seconds)
    capacity = np.bincount(cls, minlength=len(sw.classes))
    srv = S.SceneServer(sw.classes, capacity, args.serve_budget_gb * 1e9, sw.model_kwargs, sw.tips, device=dev,
                        resolution=sw.resolution, camera_pos=sw.camera_pos, look_at=sw.look_at,
                        block=args.block, streams=args.serve_streams, max_inflight=args.serve_inflight,
                        shared=args.shared, parts=args.serve_parts)
    if getattr(srv.model, "split_acc", None) is not None:
        srv.model.enable_shared(fold=args.fold == "on")
    for a in range(sw.actors):
        srv.register(a, int(cls[a]))
    log(f"serving {sw.name}: {len(reqs)} requests over {args.serve_seconds:.0f} s, "
        f"{srv.n_slots} slots, class bytes {srv.class_bytes}")
    out = S.serve(srv, reqs, args.serve_seconds)
    st = srv.stats
    out.update(workload=sw.name, duration_s=args.serve_seconds, requests=len(reqs), actors=sw.actors,
               rate_req_per_s=sw.rate, ticks=int(st["ticks"]),
               mean_tick_requests=st["tick_requests"] / max(st["ticks"], 1), admitted=int(st["admitted"]),
               streams=args.serve_streams, max_inflight=args.serve_inflight, parts=args.serve_parts,
               admission_waits=int(st["admission_waits"]), budget_gb=args.serve_budget_gb,
               resident_gb=srv.resident_bytes / 1e9,
               obs="64x64 fp32 depth + 32 scalars per request" if sw.resolution == 64 else f"{sw.resolution}^2 depth")
    del srv
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; several GPUs would each step their own "
                                   "share of the environments, which is not launched here", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.gen_workers:
        os.environ["FFSRL_GEN_WORKERS"] = str(args.gen_workers)
    wl = importlib.import_module("ff-srl_amd.workloads")
    if args.workload == "serve":
        import torch
        pkg = importlib.import_module("ff-srl_amd")
        srv = serving_run(pkg, wl, args, torch.device("cuda:0"))
        print(json.dumps({"metric": "serving: env-steps/s + p50/p99 step latency", "value": srv["env_steps_per_s"],
                          "unit": "env-steps/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
                          "data": "synthetic (seeded meshes, seeded Poisson requests)",
                          "config": {"workload": srv["workload"]}, "serving": srv}), flush=True)
        return
    t_setup = time.perf_counter()
    w = make_workload(wl, args)                       # meshes generated before the GPU runtime starts
    log(f"workload {w.name}: {w.totals()} in {time.perf_counter() - t_setup:.1f} s")
    import torch
    pkg = importlib.import_module("ff-srl_amd")
    dev = torch.device("cuda:0")
    model = wl.build_model(pkg, w, dev)
    ren = pkg.RaycastRendererDO(dev, model, cameraPos=w.camera_pos, resolution=w.resolution)
    ren.lookAt(w.look_at)
    integ = pkg.SimIntegratorDO(dev, block=args.block)
    lib = pkg.lib
    n = w.num_envs
    gen = torch.Generator(device=dev)
    gen.manual_seed(args.seed)
    total_frames = args.warmup + args.steps
    walk = torch.randn((total_frames, n, 3), generator=gen, device=dev) * w.walk_sigma
    depth_out = ren.depth
    stream = torch.cuda.current_stream()
    s = stream.cuda_stream

    cam_rgb = pkg._lib.Camera()
    C.pointer(cam_rgb)[0] = ren.camera
    cam_rgb.mode = pkg.RenderMode.texture              # texture lookup (each mesh's seeded texture), Lambert-lit

    def render(sh, eb=0, ee=n):
        if w.rgb:
            # RGB-D: normals, then colour and depth from one traversal
            lib.ffsrl_vertex_normals(model.batch, eb, ee, model.normals.data_ptr(), sh)
            lib.ffsrl_render(model.batch, cam_rgb, eb, ee, depth_out[eb:].data_ptr(), ren.pixels[eb:].data_ptr(),
                             model.normals.data_ptr(), model.vertexColors.data_ptr(), model.t_tri_colors.data_ptr(),
                             None, None, sh)
        else:
            lib.ffsrl_render(model.batch, ren.camera, eb, ee, depth_out[eb:].data_ptr(), None, None, None, None, None,
                             None, sh)

    def move(f, sh):
        """The tool random walk: instrument Cartesian actions for a model
        with laparoscopes, else a direct move of the instrument points."""
        if w.laparoscopes is not None:
            lib.ffsrl_instrument_cartesian(model.batch, model.instrument, walk[f].data_ptr(), None, None, sh)
        else:
            lib.ffsrl_move_tools(model.batch, walk[f].data_ptr(), None, None, sh)

    def step(f, sh):
        move(f, sh)
        integ.stepModel(model, refit=True)       # on torch's current stream, which is sh
        render(sh)

    pkg.SimBVH(dev, model)
    tuned = None
    if not args.no_autotune:
        t_tune = time.perf_counter()
        tuned = integ.autotune(model)
        log(f"autotuned {tuned} in {time.perf_counter() - t_tune:.1f} s")
    if args.shared != "auto":
        integ.shared = args.shared == "on"
    if args.chunk and args.no_autotune:
        model.set_constraint_order("local", args.chunk)
    if args.parts == "on":
        integ.shared, integ.parts = False, True
        model.enable_parts(tets_per_part=args.tets_per_part)
    if integ.shared:
        model.enable_shared(fold=args.fold == "on")
    if args.calibrate_copy:
        a = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t_setup
    log(f"model on device after {setup_s:.1f} s; warmup")
    for f in range(args.warmup):
        step(f, s)
    torch.cuda.synchronize()
    # every timed form (eager, graph, plain) starts from this state
    state = [t for t in (model.x, model.p, model.v, model.drag, model.laparoscopeTip, model.laparoscopeBase,
                         model.bvh_lo, model.bvh_hi, model.lam_t, model.lam_e,
                         getattr(model, "laparoscopeXForm", None)) if t is not None]
    snapshot = [(t, t.clone()) for t in state]

    def restore():
        torch.cuda.synchronize()
        for t, saved in snapshot:
            t.copy_(saved)
        torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        f = args.warmup + i
        e0, e1, e2, e3 = ev[i]
        e0.record(stream)
        move(f, s)
        e1.record(stream)
        integ.stepModel(model, refit=True)
        e2.record(stream)
        render(s)
        e3.record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {elapsed:.3f} s")
    step_ms = np.array([ev[i][0].elapsed_time(ev[i][3]) for i in range(args.steps)])
    frame_ms = np.array([ev[i][1].elapsed_time(ev[i][2]) for i in range(args.steps)])
    render_ms = np.array([ev[i][2].elapsed_time(ev[i][3]) for i in range(args.steps)])
    value = n * args.steps / elapsed
    finite = bool(torch.isfinite(model.x).all().item())
    status = model.frame_status().cpu().numpy()

    out = {
        "metric": "env-steps/s @1024 mixed-size envs + p50/p99 step latency, batched vs plain",
        "value": value, "unit": "env-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": w.data,
        "config": {"workload": w.name, "envs": n, **{k: int(v) for k, v in w.totals().items()},
                   "substeps": w.model_kwargs["simSubsteps"], "iterations": w.model_kwargs["simConstraintsSteps"],
                   "render": f"{w.resolution}x{w.resolution} depth{' + RGB (texture mode)' if w.rgb else ''} "
                             f"(2x2 supersampled)", "block": integ.block,
                   "dispatch": "lpt" if integ.dispatch_sorted else "table", "shared": integ.shared,
                   **({"fold": getattr(model, "fold_ws", None) is not None} if integ.shared else {}),
                   **({"parts": int(sum(model.parts_per_mesh[k] for k in model.env_mesh)),
                       "tets_per_part": model._parts_cfg["tets_per_part"]} if integ.parts else {}),
                   "chunk": model.topo[0].chunk, "autotuned": tuned is not None,
                   "form": ("partitioned (claim table + 1 frame kernel + 1 render kernel per step)" if integ.parts
                            else "batched (1 frame kernel + 1 render kernel per step)"),
                   **({"collision": True} if w.model_kwargs.get("collision") else {})},
        "p50_step_ms": float(np.percentile(step_ms, 50)), "p99_step_ms": float(np.percentile(step_ms, 99)),
        "frame_kernel_ms": float(frame_ms.mean()), "render_kernel_ms": float(render_ms.mean()),
        "rays_per_s": float(4 * w.resolution ** 2 * n / (render_ms.mean() * 1e-3)),
        "finite": finite, "setup_s": setup_s,
        # envs whose frames reported a saturated fixed-point term / an incomplete shared frame
        "status_envs": {"saturated": int(((status & 1) != 0).sum()), "incomplete": int(((status & 2) != 0).sum())},
    }
    alg = algorithmic_bytes(w, model)
    achieved = alg / (frame_ms.mean() * 1e-3) / 1e9
    kname = "k_frame_parts" if integ.parts else "k_frame_shared" if integ.shared else "k_frame"
    traffic, source = measured_traffic().get((w.name, kname), (None, None))
    copy_peak = copy_peak_gbs(dev)
    out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                       "kernel": kname + " (physics + refit)", "algorithmic_bytes_per_launch": alg,
                       "copy_peak": copy_peak, "frac_of_copy_peak": achieved / copy_peak}
    if traffic is not None:
        # measured HBM bytes over this run's frame time, against the nominal peak
        out["roofline"]["traffic_source"] = source
        out["roofline"]["measured_hbm_frac"] = traffic / (frame_ms.mean() * 1e-3) / 1e9 / HBM_PEAK_GBS

    if not args.no_graph:
        log("graph replay")
        restore()
        g = C.c_void_p()
        cs = torch.cuda.Stream()
        cs.wait_stream(stream)
        with torch.cuda.stream(cs):
            lib.ffsrl_graph_begin(cs.cuda_stream)
            step(total_frames - 1, cs.cuda_stream)
            lib.ffsrl_graph_end(cs.cuda_stream, C.byref(g))
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(args.steps):
                lib.ffsrl_graph_launch(g, cs.cuda_stream)
            torch.cuda.synchronize()
            out["graph_env_steps_per_s"] = n * args.steps / (time.perf_counter() - t1)
        lib.ffsrl_graph_destroy(g)

    if not args.no_plain and args.plain_frames > 0:
        log("plain form")
        restore()
        model.enable_plain()

        def plain_step(f):
            move(f % total_frames, s)
            integ.stepModelPlain(model)
            lib.ffsrl_bvh_refit_plain(model.batch, 0, n, s)
            for e in range(n):
                render(s, e, e + 1)

        plain_step(0)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        pms = []
        for f in range(args.plain_frames):
            a = time.perf_counter()
            plain_step(f)
            torch.cuda.synchronize()
            pms.append((time.perf_counter() - a) * 1e3)
        plain_el = time.perf_counter() - tp
        out["plain_env_steps_per_s"] = n * args.plain_frames / plain_el
        out["plain_p50_step_ms"] = float(np.percentile(pms, 50))
        out["batched_over_plain"] = value / out["plain_env_steps_per_s"]

        # the reference's launch model: one launch per kernel over all envs
        # (ffsrl_frame_do), then one render launch over all envs
        log("reference launch model")
        restore()

        def do_step(f):
            move(f % total_frames, s)
            lib.ffsrl_frame_do(model.batch, model.params, 0, n, 1, s)
            render(s)

        do_step(0)
        torch.cuda.synchronize()
        tp = time.perf_counter()
        dms = []
        for f in range(args.plain_frames * 5):
            a = time.perf_counter()
            do_step(f)
            torch.cuda.synchronize()
            dms.append((time.perf_counter() - a) * 1e3)
        do_el = time.perf_counter() - tp
        out["one_launch_env_steps_per_s"] = n * len(dms) / do_el
        out["one_launch_p50_step_ms"] = float(np.percentile(dms, 50))
        out["one_launch_launches_per_frame"] = one_launch_count(model, n)
        out["batched_over_one_launch"] = value / out["one_launch_env_steps_per_s"]

    if args.serve_seconds > 0:
        log("serving")
        out["serving"] = serving_run(pkg, wl, args, dev)

    if not args.no_cpu_baseline:
        log("cpu baseline")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        out["cpu_baseline"], out["cpu_baseline_1core"] = cpu_baseline(pkg, orc, w, model, ren, args.cpu_seconds,
                                                                      args.seed, args.cpu_threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
