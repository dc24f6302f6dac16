"""TEST INFRASTRUCTURE ONLY - the oracle.

CPU restatement of the reference's per-step kernels (C, ``oracle.c`` /
``oracle_kernels.h``) plus a NumPy restatement of the host BVH build
(``FF_SRL/FF_SRL/bvh.py:228-518``).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg use this module, and only as the
checker or as the timed CPU baseline; the product never imports it.

Reference layout: AoS vec3 arrays, global int32 indices (env vertex base
added), environments concatenated, kernels in serial thread order.

Parity status of each restated function (see the tests that pin them):
  gravity, PBDStep, bounds          - pinned: closed forms (tests/test_oracle.py)
  distance PBD / XPBD               - pinned: 2-particle closed forms, momentum
  volume PBD / XPBD                 - pinned: single-tet closed forms, orientation
  neoHookean                        - pinned: F = I values, finite-difference
                                      gradients, rotation invariance; the
                                      Warp matrix-constructor convention itself
                                      is parity unpinned (Warp is absent)
  apply, drag                       - pinned: closed forms
  refit                             - pinned: brute-force min/max, bit-exact
  ray cast (AABB, MT, traversal)    - pinned: brute-force nearest hit, analytic
                                      plane scene
  BVH build (bvh_build below)       - pinned: a brute-force SAH (explicit bin
                                      sets, no prefix sweep) node for node,
                                      hand-built trees, a hand-worked
                                      two-pointer partition; planted mistakes
                                      fail (tests/test_bvh_sah_pin.py); the
                                      float64 bin arithmetic is an assumption
                                      about the reference's NumPy 1.x / Warp
                                      promotion (parity unpinned at bin edges)
  instrument collision              - pinned: sphere-on-triangle closed form,
                                      no-contact cases, the barycentric-point
                                      quirk, summed responses
                                      (tests/test_collision_oracle.py)
  instrument kinematics             - pinned: float64 matrix algebra (inverse,
                                      Rx Ry Rz products), rigid-motion
                                      invariants of the actions, the
                                      remote-centre property of Cartesian
                                      actions, the jaw angle limits and clamp
                                      state table (tests/test_instrument_oracle.py);
                                      Warp's exact operation order for mat44
                                      inverse / quaternions is parity unpinned
  remap (remapAToBKernel)           - pinned: the reference's liver scene maps
                                      each render point to the simulation
                                      vertex at its position (to 1.5e-5 in a
                                      12-unit scene), the swapped map does not
                                      (tests/test_remap_oracle.py)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FFSRL_ORACLE_LIB selects another build of the same C (the sanitizer build)
LIB_PATH = os.environ.get("FFSRL_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _cfg_struct(real):
    class Cfg(C.Structure):
        _fields_ = [
            ("nV", C.c_long), ("nE", C.c_long), ("nT", C.c_long),
            ("substeps", C.c_int), ("iterations", C.c_int),
            ("dt", real), ("gravity", real * 3), ("damping", real), ("ground", real),
            ("dist_mode", C.c_int), ("vol_mode", C.c_int), ("nh_on", C.c_int), ("lambda_legacy", C.c_int),
            ("dist_k", real), ("vol_k", real), ("nh_vol_c", real), ("nh_dev_c", real), ("drag_ks", real),
            ("collide", C.c_int), ("nF", C.c_long), ("contact_k", real),
        ]
    return Cfg


CfgF32 = _cfg_struct(C.c_float)
CfgF64 = _cfg_struct(C.c_double)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class StepConfig:
    """The scalars of one simulation (SimModelDO arguments, modelDO.py:774-830)."""

    def __init__(self, substeps=10, iterations=1, fps=60.0, gravity=(0.0, -9.81, 0.0), damping=0.1, ground=0.0,
                 dist_mode=0, dist_k=0.0, vol_mode=0, vol_k=0.0, nh_on=0, nh_vol_c=1.0, nh_dev_c=1.0,
                 drag_ks=1.0, lambda_legacy=0, dist_compliance=None, vol_compliance=None, collide=0,
                 contact_k=1.0):
        self.substeps, self.iterations = substeps, iterations
        self.dt = (1.0 / fps) / substeps
        self.gravity, self.damping, self.ground = gravity, damping, ground
        self.dist_mode, self.vol_mode, self.nh_on = dist_mode, vol_mode, nh_on
        # XPBD kernels take alpha~ = compliance / dt^2 (modelDO.py:828-830).
        self.dist_k = dist_k if dist_compliance is None else dist_compliance / (self.dt * self.dt)
        self.vol_k = vol_k if vol_compliance is None else vol_compliance / (self.dt * self.dt)
        self.nh_vol_c, self.nh_dev_c = nh_vol_c, nh_dev_c
        self.drag_ks, self.lambda_legacy = drag_ks, lambda_legacy
        self.collide, self.contact_k = collide, contact_k

    def cstruct(self, dtype, nV, nE, nT, nF=0):
        S = CfgF32 if dtype == np.float32 else CfgF64
        c = S()
        c.nV, c.nE, c.nT = nV, nE, nT
        c.substeps, c.iterations = self.substeps, self.iterations
        c.dt = self.dt
        c.gravity[:] = self.gravity
        c.damping, c.ground = self.damping, self.ground
        c.dist_mode, c.vol_mode, c.nh_on = self.dist_mode, self.vol_mode, self.nh_on
        c.lambda_legacy = self.lambda_legacy
        c.dist_k, c.vol_k = self.dist_k, self.vol_k
        c.nh_vol_c, c.nh_dev_c = self.nh_vol_c, self.nh_dev_c
        c.drag_ks = self.drag_ks
        c.collide, c.nF, c.contact_k = int(self.collide), nF, self.contact_k
        return c


class OracleSim:
    """Concatenated multi-env state in the reference layout."""

    def __init__(self, meshes, env_mesh, dtype=np.float32, initial_velocity=None, tips=None, caps_base=None,
                 caps_radius=None):
        self.dtype = dtype
        vb = 0
        edges, rest, tets, vol, inv, x, w, v2e, tri, t2e = [], [], [], [], [], [], [], [], [], []
        self.vbase = [0]
        for i, k in enumerate(env_mesh):
            m = meshes[k]
            edges.append(m.edges.astype(np.int32) + vb)
            rest.append(m.edge_rest)
            tets.append(m.tets.astype(np.int32) + vb)
            vol.append(m.tet_rest_volume)
            inv.append(m.tet_inv_rest.reshape(-1, 9))
            x.append(m.verts)
            w.append(m.inv_mass)
            v2e.append(np.full(m.num_vertices, i, np.int32))
            tri.append(m.tris.astype(np.int32) + vb)
            t2e.append(np.full(m.num_tris, i, np.int32))
            vb += m.num_vertices
            self.vbase.append(vb)
        cat = np.concatenate
        self.nV, self.nE, self.nT = vb, sum(len(e) for e in edges), sum(len(t) for t in tets)
        self.edge = np.ascontiguousarray(cat(edges).reshape(-1)).astype(np.int32)
        self.tet = np.ascontiguousarray(cat(tets).reshape(-1)).astype(np.int32)
        self.edge_rest = cat(rest).astype(dtype)
        self.rest_vol = cat(vol).astype(dtype)
        self.inv_rest = np.ascontiguousarray(cat(inv)).astype(dtype)
        self.x = np.ascontiguousarray(cat(x)).astype(dtype)
        self.inv_mass = cat(w).astype(dtype)
        self.v = np.zeros_like(self.x) if initial_velocity is None else np.ascontiguousarray(
            np.asarray(initial_velocity).reshape(-1, 3)).astype(dtype)
        self.p = self.x.copy()
        self.dP = np.zeros_like(self.x)
        self.cnt = np.zeros(self.nV, np.int32)
        self.lam_e = np.zeros(max(self.nE, 1), dtype)
        self.lam_t = np.zeros(max(self.nT, 1), dtype)
        self.drag = np.zeros(self.nV, dtype)
        self.v2e = cat(v2e).astype(np.int32)
        self.tri = np.ascontiguousarray(cat(tri).reshape(-1)).astype(np.int32)
        self.t2e = cat(t2e).astype(np.int32)
        self.nF = len(self.t2e)
        n_env = len(env_mesh)
        self.tips = np.zeros((n_env * 4, 3), dtype) if tips is None else np.ascontiguousarray(
            np.asarray(tips).reshape(-1, 3)).astype(dtype)
        self.caps_base = np.zeros((n_env * 4, 3), dtype) if caps_base is None else np.ascontiguousarray(
            np.asarray(caps_base).reshape(-1, 3)).astype(dtype)
        self.caps_radius = np.zeros(n_env * 4, dtype) if caps_radius is None else np.ascontiguousarray(
            np.asarray(caps_radius).reshape(-1)).astype(dtype)
        self.in_collision = np.zeros(n_env, np.int32)

    def step(self, cfg: StepConfig, frames: int = 1, use_drag: bool = True):
        L = lib()
        sfx = "_f32" if self.dtype == np.float32 else "_f64"
        fn = getattr(L, "oracle_step_frame" + sfx)
        c = cfg.cstruct(self.dtype, self.nV, self.nE, self.nT, self.nF)
        for _ in range(frames):
            fn(C.byref(c), _p(self.x), _p(self.p), _p(self.v), _p(self.inv_mass), _p(self.dP), _p(self.cnt),
               _p(self.edge), _p(self.edge_rest), _p(self.tet), _p(self.rest_vol), _p(self.inv_rest),
               _p(self.lam_e), _p(self.lam_t), _p(self.drag) if use_drag else None, _p(self.v2e), _p(self.tips),
               _p(self.tri), _p(self.t2e), _p(self.caps_base), _p(self.caps_radius), _p(self.in_collision))


def call(name, dtype, *args):
    sfx = "_f32" if dtype == np.float32 else "_f64"
    fn = getattr(lib(), name + sfx)
    conv = []
    real = C.c_float if dtype == np.float32 else C.c_double
    for a in args:
        if isinstance(a, np.ndarray):
            conv.append(C.c_void_p(a.ctypes.data))
        elif isinstance(a, float):
            conv.append(real(a))
        elif a is None:
            conv.append(None)
        else:
            conv.append(C.c_long(a))
    fn(*conv)


def refit(bvh, faces, vertex, dtype=np.float32):
    """Restated refitBVHLevel loop for one env (bvh.py:30-104, 419-440)."""
    n = len(bvh["left"])
    mn = np.zeros((n, 3), dtype)
    mx = np.zeros((n, 3), dtype)
    sfx = "_f32" if dtype == np.float32 else "_f64"
    fn = getattr(lib(), "oracle_refit" + sfx)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    arrs = [i32(bvh["left"]), i32(bvh["first"]), i32(bvh["count"]), i32(bvh["level"]), i32(bvh["tri_map"]),
            i32(np.asarray(faces).reshape(-1))]
    vert = np.ascontiguousarray(vertex, dtype=dtype)
    fn(C.c_int(n), C.c_int(int(np.max(bvh["level"]))), *[_p(a) for a in arrs], _p(vert), _p(mn), _p(mx))
    return mn, mx


def render_depth(H, W, origin, dirs, bvh, aabb_min, aabb_max, faces, vertex, dtype=np.float32):
    """Restated drawSimBVHKernelManyEnvs (depth mode) + downsample for one env.
    Returns depth (H, W), per-ray hit face (2H*2W) and t, and the deepest stack use."""
    depth = np.zeros((H, W), dtype)
    hit = np.zeros(4 * H * W, np.int32)
    t = np.zeros(4 * H * W, dtype)
    ms = C.c_int(0)
    sfx = "_f32" if dtype == np.float32 else "_f64"
    fn = getattr(lib(), "oracle_render_depth" + sfx)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f = lambda a: np.ascontiguousarray(a, dtype=dtype)
    arrs = [f(origin), f(np.asarray(dirs)[:, :3]), f(aabb_min), f(aabb_max), i32(bvh["left"]), i32(bvh["first"]),
            i32(bvh["count"]), i32(bvh["tri_map"]), i32(np.asarray(faces).reshape(-1)), f(vertex)]
    fn(C.c_int(H), C.c_int(W), *[_p(a) for a in arrs], _p(depth), _p(hit), _p(t), C.byref(ms))
    return depth, hit, t, ms.value


def vertex_normals(faces, vertex, dtype=np.float32):
    """Restated zeroVec3Array + vertex_normal_sum_kernel + normalize_kernel
    (render.py:81-121): area-weighted normals summed in face order."""
    faces = np.ascontiguousarray(np.asarray(faces).reshape(-1, 3), dtype=np.int32)
    vert = np.ascontiguousarray(vertex, dtype=dtype)
    out = np.zeros_like(vert)
    sfx = "_f32" if dtype == np.float32 else "_f64"
    getattr(lib(), "oracle_vertex_normals" + sfx)(C.c_long(faces.shape[0]), _p(faces), C.c_long(vert.shape[0]),
                                                   _p(vert), _p(out))
    return out


def render_rgb(H, W, mode, origin, dirs, bvh, aabb_min, aabb_max, faces, vertex, normals, vcolor=None, tcolor=None,
               light=(3.0, 0.0, 0.0), intensity=0.3, dtype=np.float32, tri_uv=None, tri_tex=None, texels=None,
               tex_info=None):
    """Restated drawSimBVHKernelManyEnvs colour modes (0 grayscale, 1 texture
    with the vertex-colour fallback, 2 normal map, 3 vertex colour, 5
    per-face colour) + per-channel downsample.  Texture inputs, by original
    face id: tri_uv (F, 3, 2) corner texture coordinates, tri_tex (F,) texture
    id or -1; texels (N, 3) of every texture; tex_info (n_tex, 2) = (first
    texel, side).  Returns rgb (H, W, 3)."""
    rgb = np.zeros((H, W, 3), dtype)
    sfx = "_f32" if dtype == np.float32 else "_f64"
    fn = getattr(lib(), "oracle_render_rgb" + sfx)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=dtype)
    nf = np.asarray(faces).reshape(-1, 3).shape[0]
    if tri_tex is None:
        tri_tex = np.full(nf, -1, np.int32)
    if tri_uv is None:
        tri_uv = np.zeros((nf, 3, 2))
    texels = np.zeros((1, 3)) if texels is None else np.asarray(texels).reshape(-1, 3)
    tex_info = np.zeros((1, 2), np.int64) if tex_info is None else np.asarray(tex_info, np.int64).reshape(-1, 2)
    arrs = [f(origin), f(np.asarray(dirs)[:, :3]), f(aabb_min), f(aabb_max), i32(bvh["left"]), i32(bvh["first"]),
            i32(bvh["count"]), i32(bvh["tri_map"]), i32(np.asarray(faces).reshape(-1)), f(vertex), f(normals),
            f(vcolor), f(tcolor), f(np.asarray(tri_uv).reshape(-1)), i32(tri_tex), f(texels)]
    info = np.ascontiguousarray(tex_info, np.int64)
    fn(C.c_int(H), C.c_int(W), C.c_int(mode), *[_p(a) for a in arrs], C.c_long(texels.shape[0]), _p(info),
       _p(f(np.asarray(light))), (C.c_float if dtype == np.float32 else C.c_double)(intensity), _p(rgb))
    return rgb


def texture_lookup(s, t, texels, shift, size, dtype=np.float32):
    """Restated texture_interpolation (render.py:124-148) at texture
    coordinates (s, t) of the square texture of side `size` starting at
    texel `shift` of the table `texels` (N, 3)."""
    texels = np.ascontiguousarray(np.asarray(texels).reshape(-1, 3), dtype)
    out = np.zeros(3, dtype)
    real = C.c_float if dtype == np.float32 else C.c_double
    sfx = "_f32" if dtype == np.float32 else "_f64"
    getattr(lib(), "oracle_texture_lookup" + sfx)(real(s), real(t), _p(texels), C.c_long(texels.shape[0]),
                                                  C.c_long(shift), C.c_int(size), _p(out))
    return out


# ------------------------------------------------------------ BVH build
MUTANTS = ("partition_le", "right_sweep_off_by_one", "no_bin_clamp")


def sah_partition(tri, first, count, cen_axis, pos, le=False):
    """The in-place two-pointer partition of subdivide (bvh.py:367-376):
    tri[first:first+count] is reordered so that the triangles whose centroid
    coordinate is < pos come first; returns their number.  ``le`` (test
    mutant only) takes <= instead of <."""
    i, j = first, first + count - 1
    while i <= j:
        c = float(cen_axis[tri[i]])
        if (c <= pos) if le else (c < pos):
            i += 1
        else:
            tri[i], tri[j] = tri[j], tri[i]
            j -= 1
    return i - first


def bvh_build(vertex, faces, bins=8, mutant=None):
    """NumPy restatement of SimBVH.buildBVH (bvh.py:456-518) with
    findBestSplitPlaneBinned (bvh.py:281-340) and subdivide (bvh.py:354-401).

    Centroids are float32 (calculateCentroids, bvh.py:6-21); bin positions,
    areas and costs are float64.  The float64 choice is an assumption about
    the reference's promotion: it mixes np.float32 centroids with Python
    floats (and the Python-scope wp.min of Warp 0.10.1) in
    findBestSplitPlaneBinned; the min/max centroid bounds are exact either
    way, and a float32 evaluation could bin differently only a centroid
    within rounding of a bin edge.  Returns left / first / count / level
    arrays (nodes in allocation order), the triangle permutation and the max
    level.

    ``mutant`` plants one of MUTANTS (a plausible misreading of the
    reference) so that tests can show their checks catch it; it is never set
    outside tests.
    """
    if mutant is not None and mutant not in MUTANTS:
        raise ValueError(f"unknown mutant {mutant!r}")
    vertex = np.asarray(vertex, dtype=np.float32)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    nf = faces.shape[0]
    tv = vertex[faces]                                            # (F, 3, 3) float32
    cen = ((tv[:, 0] + tv[:, 1]) + tv[:, 2]) / np.float32(3.0)    # float32
    tv64 = tv.astype(np.float64)
    tmin, tmax = tv64.min(axis=1), tv64.max(axis=1)               # per-triangle bounds
    tri = np.arange(nf, dtype=np.int64)
    cap = 2 * nf
    left = np.full(cap, -1, np.int64)
    first = np.zeros(cap, np.int64)
    count = np.zeros(cap, np.int64)
    level = np.full(cap, -1, np.int64)
    bmin = np.zeros((cap, 3))
    bmax = np.zeros((cap, 3))
    first[0], count[0], level[0] = 0, nf, 0
    bmin[0] = vertex.min(axis=0).astype(np.float64)
    bmax[0] = vertex.max(axis=0).astype(np.float64)
    used = [1]

    def area(lo, hi):
        e = hi - lo
        return e[..., 0] * e[..., 1] + e[..., 1] * e[..., 2] + e[..., 2] * e[..., 0]

    def best_split(node):
        ids = tri[first[node]:first[node] + count[node]]
        best, best_axis, best_pos = 1e30, -1, 0.0
        for axis in range(3):
            c = cen[ids, axis].astype(np.float64)
            lo_c, hi_c = c.min(), c.max()
            if hi_c == lo_c:
                continue
            scale = bins / (hi_c - lo_c)
            b = ((c - lo_c) * scale).astype(np.int64)
            if mutant == "no_bin_clamp":        # the top centroid lands past the last bin
                keep = b < bins
                b, bt_lo, bt_hi = b[keep], tmin[ids][keep], tmax[ids][keep]
            else:
                b = np.minimum(bins - 1, b)
                bt_lo, bt_hi = tmin[ids], tmax[ids]
            cnt = np.bincount(b, minlength=bins)
            blo = np.full((bins, 3), 1e30)
            bhi = np.full((bins, 3), -1e30)
            np.minimum.at(blo, b, bt_lo)
            np.maximum.at(bhi, b, bt_hi)
            lcnt = np.cumsum(cnt)[:-1]
            rcnt = np.cumsum(cnt[::-1])[:-1][::-1]
            llo = np.minimum.accumulate(blo, axis=0)[:-1]
            lhi = np.maximum.accumulate(bhi, axis=0)[:-1]
            rlo = np.minimum.accumulate(blo[::-1], axis=0)[:-1][::-1]
            rhi = np.maximum.accumulate(bhi[::-1], axis=0)[:-1][::-1]
            if mutant == "right_sweep_off_by_one":   # right side of plane j starts one bin late
                rcnt = np.append(rcnt[1:], 0)
                rlo = np.vstack([rlo[1:], np.full((1, 3), 1e30)])
                rhi = np.vstack([rhi[1:], np.full((1, 3), -1e30)])
            cost = lcnt * area(llo, lhi) + rcnt * area(rlo, rhi)
            step = (hi_c - lo_c) / bins
            for j in range(bins - 1):
                if cost[j] < best:
                    best_axis, best_pos, best = axis, lo_c + step * (j + 1), cost[j]
        return best, best_axis, best_pos

    def subdivide(node, lvl):
        best, axis, pos = best_split(node)
        parent = count[node] * area(bmin[node], bmax[node])
        if best >= parent:
            return
        lc = sah_partition(tri, first[node], count[node], cen[:, axis], pos, le=(mutant == "partition_le"))
        i = first[node] + lc
        if lc == 0 or lc == count[node]:
            return
        a, b = used[0], used[0] + 1
        used[0] += 2
        first[a], count[a], level[a] = first[node], lc, lvl + 1
        first[b], count[b], level[b] = i, count[node] - lc, lvl + 1
        left[node] = a
        count[node] = 0
        for ch in (a, b):
            ids = tri[first[ch]:first[ch] + count[ch]]
            bmin[ch] = tmin[ids].min(axis=0)
            bmax[ch] = tmax[ids].max(axis=0)
        subdivide(a, lvl + 1)
        subdivide(b, lvl + 1)

    subdivide(0, 0)
    n = used[0]
    return {"left": left[:n].astype(np.int32), "first": first[:n].astype(np.int32),
            "count": count[:n].astype(np.int32), "level": level[:n].astype(np.int32),
            "tri_map": tri.astype(np.int32), "max_level": int(level[:n].max())}


def remap(dst, src, pairs):
    """remapAToBKernel (``FF_SRL/FF_SRL/utils.py:32-42``): for each pair
    (point, vertex), in thread order, ``dst[point] = src[vertex]``; a point
    listed twice keeps its last pair's value (the scene's maps list each
    point once).  Returns a new array; dst is not modified."""
    out = np.array(dst, copy=True)
    pairs = np.asarray(pairs).reshape(-1, 2)
    for pt, vx in pairs:
        out[pt] = src[vx]
    return out


def collide(pred, tri, tri_env, inv_mass, caps_base, caps_tip, caps_radius, k=1.0, dtype=np.float64):
    """Restated collideLaparoscopeDO (integrator.py:981-1050) for one pass.
    Returns the updated positions and the per-env triangle-in-collision count."""
    pred = np.ascontiguousarray(pred, dtype).copy()
    tri = np.ascontiguousarray(np.asarray(tri).reshape(-1), np.int32)
    tri_env = np.ascontiguousarray(tri_env, np.int32)
    n_env = int(np.asarray(caps_radius).size // 4)
    inc = np.zeros(n_env, np.int32)
    acc = np.zeros_like(pred)
    f = lambda a: np.ascontiguousarray(a, dtype)
    real = C.c_float if dtype == np.float32 else C.c_double
    sfx = "_f32" if dtype == np.float32 else "_f64"
    getattr(lib(), "oracle_collide" + sfx)(C.c_long(tri.size // 3), _p(pred), _p(tri), _p(tri_env), _p(f(inv_mass)),
                                           _p(f(caps_base)), _p(f(caps_tip)), _p(f(caps_radius)), _p(inc), real(k),
                                           C.c_long(pred.shape[0]), _p(acc))
    return pred, inc


class OracleInstrument:
    """Restated instrument kinematics of n envs (modelDO.py:55-289, 1331-1404,
    1570-1587): seven transforms per env stored transposed (``xf``, n*7
    matrices), capsule heights (n, 4), the clamp state ``dch`` / ``dchu``
    (n*3 int32, drag first), workspace ``violation`` (n int32) and the
    capsule points ``tips`` / ``bases`` (n*4, 3)."""

    def __init__(self, xf, height, tips, bases, dtype=np.float32):
        self.dtype = dtype
        self.xf = np.ascontiguousarray(np.asarray(xf).reshape(-1, 4, 4), dtype).copy()
        self.n = self.xf.shape[0] // 7
        self.height = np.ascontiguousarray(np.asarray(height).reshape(-1), dtype).copy()
        self.tips = np.ascontiguousarray(np.asarray(tips).reshape(-1, 3), dtype).copy()
        self.bases = np.ascontiguousarray(np.asarray(bases).reshape(-1, 3), dtype).copy()
        self.dch = np.zeros(3 * self.n, np.int32)
        self.dchu = np.zeros(3 * self.n, np.int32)
        self.violation = np.zeros(self.n, np.int32)

    def _fn(self, name):
        return getattr(lib(), name + ("_f32" if self.dtype == np.float32 else "_f64"))

    def _f(self, a):
        return np.ascontiguousarray(np.asarray(a).reshape(-1), self.dtype)

    def actions(self, actions):
        """applyActions without the grasp selection: (n, 5) actions."""
        a = self._f(actions)
        self._fn("oracle_instrument_actions")(C.c_int(self.n), _p(self.xf), _p(self.height), _p(self.dch),
                                              _p(self.dchu), _p(a), _p(self.tips), _p(self.bases))

    def cartesian(self, actions, lo=None, hi=None):
        a = self._f(actions)
        lo = None if lo is None else self._f(lo)
        hi = None if hi is None else self._f(hi)
        self._fn("oracle_instrument_cartesian")(C.c_int(self.n), _p(self.xf), _p(self.height), _p(a),
                                                None if lo is None else _p(lo), None if hi is None else _p(hi),
                                                _p(self.violation), _p(self.tips), _p(self.bases))

    def set_rcm(self, translation, rotation):
        t, r = self._f(translation), self._f(rotation)
        self._fn("oracle_instrument_set_rcm")(C.c_int(self.n), _p(self.xf), _p(self.height), _p(t), _p(r),
                                              _p(self.tips), _p(self.bases))

    def update(self):
        self._fn("oracle_instrument_update")(C.c_int(self.n), _p(self.xf), _p(self.height), _p(self.tips),
                                             _p(self.bases))

    def matrix(self, env, k):
        """Column-vector form of transform k of env."""
        return self.xf[env * 7 + k].T.copy()


def m4_inverse(m, dtype=np.float64):
    m = np.ascontiguousarray(m, dtype)
    out = np.zeros((4, 4), dtype)
    sfx = "_f32" if dtype == np.float32 else "_f64"
    getattr(lib(), "oracle_m4_inverse" + sfx)(_p(m), _p(out))
    return out


def m4_transformation(sx, sy, sz, ax, ay, az, dtype=np.float64):
    """getTransformationMatrix (modelDO.py:396-416), column-vector form."""
    real = C.c_float if dtype == np.float32 else C.c_double
    out = np.zeros((4, 4), dtype)
    sfx = "_f32" if dtype == np.float32 else "_f64"
    getattr(lib(), "oracle_m4_transformation" + sfx)(*(real(v) for v in (sx, sy, sz, ax, ay, az)), _p(out))
    return out
