"""The C ABI: the library loads, exports every function include/ffsrl.h
declares, the ctypes structures match the C layout, and argument errors come
back as error codes with a message (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ffsrl.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(ffsrl_\w+)\s*\(", text, flags=re.M)))


def test_every_declared_symbol_is_exported(pkg):
    h = pkg.lib.load()
    names = _declared()
    assert len(names) >= 20
    nm = subprocess.run(["nm", "-D", "--defined-only", pkg._lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ffsrl_\w+)", nm))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(h, n)
    assert set(pkg._lib.exported_symbols()) == set(names)


def test_abi_version(pkg):
    assert pkg.lib.ffsrl_abi_version() == 14
    assert pkg.lib.ffsrl_frame_max_vertices() >= 10000


_LAYOUT_C = r"""
#include <stddef.h>
#include <stdio.h>
#include "ffsrl.h"
#define F(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("ffsrl_env_desc size %zu\n", sizeof(ffsrl_env_desc));
  printf("ffsrl_batch size %zu\n", sizeof(ffsrl_batch));
  printf("ffsrl_params size %zu\n", sizeof(ffsrl_params));
  printf("ffsrl_camera size %zu\n", sizeof(ffsrl_camera));
  printf("ffsrl_requests size %zu\n", sizeof(ffsrl_requests));
  printf("ffsrl_instrument size %zu\n", sizeof(ffsrl_instrument));
  printf("ffsrl_bvh_job size %zu\n", sizeof(ffsrl_bvh_job));
  printf("ffsrl_parts size %zu\n", sizeof(ffsrl_parts));
  %FIELDS%
  return 0;
}
"""


def test_struct_layouts_match_header(pkg, tmp_path):
    L = pkg._lib
    structs = {"ffsrl_batch": L.Batch, "ffsrl_params": L.Params, "ffsrl_camera": L.Camera,
               "ffsrl_env_desc": L.EnvDesc, "ffsrl_requests": L.Requests, "ffsrl_instrument": L.Instrument,
               "ffsrl_bvh_job": L.BvhJob, "ffsrl_parts": L.Parts}
    fields = "\n".join(f"F({n}, {f[0]})" for n, S in structs.items() for f in S._fields_)
    src = tmp_path / "layout.c"
    src.write_text(_LAYOUT_C.replace("%FIELDS%", fields))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            a, b, c = line.split()
            got[(a, b)] = int(c)
    for n, S in structs.items():
        assert got[(n, "size")] == C.sizeof(S), n
        for f in S._fields_:
            assert got[(n, f[0])] == getattr(S, f[0]).offset, (n, f[0])


def test_argument_errors_are_reported(pkg):
    L = pkg._lib
    b = L.Batch()
    prm = L.Params()
    with pytest.raises(pkg.FFSRLError, match="NULL"):
        pkg.lib.ffsrl_frame(b, prm, None, 1024, 0, None)
    with pytest.raises(pkg.FFSRLError, match="params NULL"):
        pkg.lib.ffsrl_frame(b, None, None, 1024, 0, None)
    cam = L.Camera()
    cam.width = cam.height = 6
    with pytest.raises(pkg.FFSRLError):
        pkg.lib.ffsrl_render(b, cam, 0, 0, None, None, None, None, None, None, None, None)
    with pytest.raises(pkg.FFSRLError, match="remap"):
        pkg.lib.ffsrl_remap(None, None, None, 5, None)
    assert "remap" in pkg.lib.ffsrl_last_error().decode()
    # list forms of the serving loop
    dummy = np.zeros(16, np.float32)
    for f in ("envs", "x", "p", "v", "drag", "tips"):
        setattr(b, f, dummy.ctypes.data)
    b.n_env = 4
    prm.substeps = prm.iterations = 1
    prm.dt = 0.01
    with pytest.raises(pkg.FFSRLError, match="env list NULL"):
        pkg.lib.ffsrl_frame_envs(b, prm, None, 2, 1024, 0, None)
    with pytest.raises(pkg.FFSRLError, match="5 environments listed"):
        pkg.lib.ffsrl_frame_envs(b, prm, dummy.ctypes.data, 5, 1024, 0, None)
    with pytest.raises(pkg.FFSRLError, match="env list"):
        pkg.lib.ffsrl_render_envs(b, cam, None, 2, None, None, None, None, None, None, None, None)
    r = L.Requests()
    r.n = 2
    with pytest.raises(pkg.FFSRLError, match="NULL request array"):
        pkg.lib.ffsrl_apply_requests(b, r, 0.1, None)
    r.n = 9
    with pytest.raises(pkg.FFSRLError, match="9 requests"):
        pkg.lib.ffsrl_apply_requests(b, r, 0.1, None)
    with pytest.raises(pkg.FFSRLError, match="NULL argument"):
        pkg.lib.ffsrl_observe_envs(b, None, 2, None, 0.0, None, None)
    # the shared-env frame checks its scratch and counts before any device call
    b.status = dummy.ctypes.data
    with pytest.raises(pkg.FFSRLError, match="split_acc and split_ctr"):
        pkg.lib.ffsrl_frame_shared(b, prm, dummy.ctypes.data, 2, 1024, 0, None)
    with pytest.raises(pkg.FFSRLError, match="5 environments listed"):
        pkg.lib.ffsrl_frame_shared(b, prm, dummy.ctypes.data, 5, 1024, 0, None)
    b.split_acc = b.split_ctr = dummy.ctypes.data
    b.status = None
    with pytest.raises(pkg.FFSRLError, match="status array"):
        pkg.lib.ffsrl_frame_shared(b, prm, None, 0, 1024, 0, None)
    b.status = dummy.ctypes.data
    assert pkg.lib.ffsrl_frame_shared(b, prm, None, 0, 1024, 0, None) == 0      # nothing listed: no launch
    # the render's traversal choice is checked before the empty range returns
    for f in ("tris", "bvh_lo", "bvh_hi"):
        setattr(b, f, dummy.ctypes.data)
    cam = L.Camera()
    cam.dirs = dummy.ctypes.data
    cam.width = cam.height = 8
    cam.mode = 4
    cam.traversal = 3
    with pytest.raises(pkg.FFSRLError, match="render traversal 3"):
        pkg.lib.ffsrl_render(b, cam, 0, 0, dummy.ctypes.data, None, None, None, None, None, None, None)
    for t in (0, 1, 2):
        cam.traversal = t
        assert pkg.lib.ffsrl_render(b, cam, 0, 0, dummy.ctypes.data, None, None, None, None, None, None, None) == 0
    # the partitioned frame checks its tables before any device call
    pt = L.Parts()
    with pytest.raises(pkg.FFSRLError, match="parts NULL"):
        pkg.lib.ffsrl_frame_parts(b, prm, None, None, 0, 0, None)
    with pytest.raises(pkg.FFSRLError, match="NULL table"):
        pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None)
    for f in ("parts", "env_parts", "verts", "sums", "ctr", "claims", "host_env_parts"):
        setattr(pt, f, dummy.ctypes.data)
    pt.max_local = pkg.lib.ffsrl_parts_max_local() + 1
    with pytest.raises(pkg.FFSRLError, match="local vertices"):
        pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None)
    pt.max_local, pt.max_solo = 100, pkg.lib.ffsrl_parts_max_solo() + 1
    with pytest.raises(pkg.FFSRLError, match="in one part"):
        pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None)
    pt.max_solo = 0
    # the collision pass needs the batch's collision arrays and the parts' triangles
    prm.collide = 1
    for f in ("tris", "caps_base", "caps_radius", "in_collision"):
        setattr(b, f, dummy.ctypes.data)
    with pytest.raises(pkg.FFSRLError, match="parts' triangles"):
        pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None)
    pt.tris16 = dummy.ctypes.data
    assert pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None) == 0
    for f in ("tris", "caps_base", "caps_radius", "in_collision"):
        setattr(b, f, None)
    with pytest.raises(pkg.FFSRLError, match="in_collision"):
        pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None)
    prm.collide = 0
    with pytest.raises(pkg.FFSRLError, match="5 environments listed"):
        pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 5, 0, None)
    assert pkg.lib.ffsrl_frame_parts(b, prm, pt, None, 0, 0, None) == 0      # nothing listed: no launch
    assert 2000 <= pkg.lib.ffsrl_parts_max_local() < 3000
    assert pkg.lib.ffsrl_parts_max_solo() == 5088


def test_frame_rejects_too_many_terms_per_vertex(pkg):
    """The batched frames' fixed-point sums hold FFSRL_MAX_TERMS terms per
    vertex and pass; an env above it (from the host mirror of the env table)
    is rejected before any device call."""
    L = pkg._lib
    b = L.Batch()
    dummy = np.zeros(64, np.float32)
    for f in ("envs", "x", "p", "v", "drag", "tips", "status"):
        setattr(b, f, dummy.ctypes.data)
    envs = (L.EnvDesc * 2)()
    envs[1].max_terms = L.MAX_TERMS + 1
    b.host_envs = C.addressof(envs)
    b.n_env = 2
    prm = L.Params()
    prm.substeps = prm.iterations = 1
    prm.dt = 0.01
    with pytest.raises(pkg.FFSRLError, match="env 1: a vertex receives 129 Jacobi terms"):
        pkg.lib.ffsrl_frame(b, prm, None, 1024, 0, None)
    envs[1].max_terms = L.MAX_TERMS
    assert pkg.lib.ffsrl_frame_envs(b, prm, dummy.ctypes.data, 0, 1024, 0, None) == 0   # nothing listed


def test_model_counts_terms_per_vertex(pkg):
    """max_terms of the env table: incident tets plus edges per vertex, or
    incident surface triangles with collision, whichever is larger."""
    m = pkg.scenes.cube_grid(5)
    model = pkg.SimModelDO([m], [0], device="cpu", distance="xpbd", volume="xpbd")
    inc = np.bincount(m.tets.reshape(-1), minlength=m.num_vertices) + \
        np.bincount(m.edges.reshape(-1), minlength=m.num_vertices)
    F = pkg._lib.ENV_DESC_FIELDS.index("max_terms")
    assert model.env_desc_host[0, F] == inc.max()
    only_edges = pkg.SimModelDO([m], [0], device="cpu", distance="xpbd", volume=None)
    assert only_edges.env_desc_host[0, F] == np.bincount(m.edges.reshape(-1)).max()


def test_frame_rejects_envs_beyond_lds(pkg):
    L = pkg._lib
    b = L.Batch()
    dummy = np.zeros(16, np.float32)
    for f in ("envs", "x", "p", "v", "drag", "tips"):
        setattr(b, f, dummy.ctypes.data)
    b.max_nv = pkg.lib.ffsrl_frame_max_vertices() + 2
    prm = L.Params()
    prm.substeps = prm.iterations = 1
    prm.dt = 0.01
    with pytest.raises(pkg.FFSRLError, match="LDS"):
        pkg.lib.ffsrl_frame(b, prm, None, 1024, 0, None)


def test_header_has_no_hip_types():
    text = open(HEADER).read()
    assert "hipStream_t" not in re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    assert "#include <hip" not in text


def test_pack16_layout(pkg):
    """Packed constraint ids of the batched frame: 16-bit fields, first id in
    the low bits (the layout ffsrl_batch.tets16 / edges16 declares)."""
    t = np.array([[1, 2, 3, 4], [65535, 0, 10239, 7]], np.int32)
    p = pkg.model._pack16(t).view(np.uint64)
    for row, w in zip(t, p):
        assert [(int(w) >> (16 * k)) & 0xffff for k in range(4)] == list(row)
    e = np.array([[5, 9], [10239, 0]], np.int32)
    q = pkg.model._pack16(e).view(np.uint32)
    assert [(int(q[0]) & 0xffff, int(q[0]) >> 16), (int(q[1]) & 0xffff, int(q[1]) >> 16)] == [(5, 9), (10239, 0)]


def test_noalloc_frame_limits(pkg):
    """ffsrl_frame_shared_noalloc refuses batches beyond its fixed workspace
    (list slots of max_nv vertices each, 4096 envs) before any device call."""
    L = pkg._lib
    b = L.Batch()
    dummy = np.zeros(64, np.float32)
    for f in ("envs", "x", "p", "v", "drag", "tips", "status"):
        setattr(b, f, dummy.ctypes.data)
    b.n_env, b.max_nv = 5000, 10000
    prm = L.Params()
    prm.substeps = prm.iterations = 1
    prm.dt = 0.01
    with pytest.raises(pkg.FFSRLError, match="300 envs x 10000 vertices exceed its workspace"):
        pkg.lib.ffsrl_frame_shared_noalloc(b, prm, dummy.ctypes.data, 300, 1024, 0, None)
    b.max_nv = 1
    with pytest.raises(pkg.FFSRLError, match="4097 envs"):
        pkg.lib.ffsrl_frame_shared_noalloc(b, prm, dummy.ctypes.data, 4097, 1024, 0, None)
    assert pkg.lib.ffsrl_frame_shared_noalloc(b, prm, None, 0, 1024, 0, None) == 0      # nothing listed
