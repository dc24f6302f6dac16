"""ctypes binding of the C ABI declared in include/ffsrl.h.

The library is the only compute path: if ``_build/libffsrl.so`` is missing
this module raises on first use; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FFSRL_LIB selects another build of the same library (e.g. a variant built
# with different tuning flags); the default is the in-tree build.
LIB_PATH = os.environ.get("FFSRL_LIB") or os.path.join(HERE, "_build", "libffsrl.so")

OFF, PBD, XPBD = 0, 1, 2
E_ARG, E_LIMIT, E_STATE = -1, -2, -3
ABI_VERSION = 14
STATUS_SATURATED, STATUS_INCOMPLETE = 1, 2   # FFSRL_STATUS_* bits of ffsrl_batch.status
MAX_TERMS = 128                              # FFSRL_MAX_TERMS
PART_OWNER, PART_SHARED = 0x80000000, 0x40000000  # FFSRL_PART_* flags of ffsrl_parts.verts


class EnvDesc(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "vbase", "nv", "ebase", "ne", "tbase", "nt", "fbase", "nf", "nodebase", "nn",
        "boundbase", "lvlbase", "nlvl", "lam_ebase", "lam_tbase", "dp_exp", "vtopo_base", "bvh_depth",
        "max_terms", "reserved2")]


ENV_DESC_FIELDS = [f[0] for f in EnvDesc._fields_]
assert C.sizeof(EnvDesc) == 80

_vp = C.c_void_p


class Batch(C.Structure):
    _fields_ = [
        ("n_env", C.c_int32), ("max_nv", C.c_int32), ("n_tet_total", C.c_int64),
        ("envs", _vp),
        ("edges", _vp), ("edge_rest", _vp), ("tets", _vp), ("tets16", _vp), ("edges16", _vp), ("tet_rest_volume", _vp),
        ("tet_inv_rest", _vp), ("tris", _vp), ("node_left", _vp), ("node_first", _vp),
        ("node_count", _vp), ("level_nodes", _vp), ("level_info", _vp), ("level_off", _vp), ("vf_off", _vp),
        ("vf_idx", _vp), ("tri_uv", _vp), ("tri_tex", _vp), ("texels", _vp), ("tex_info", _vp),
        ("n_texels", C.c_int64),
        ("x", _vp), ("p", _vp), ("v", _vp), ("drag", _vp), ("lam_e", _vp), ("lam_t", _vp),
        ("bvh_lo", _vp), ("bvh_hi", _vp), ("tips", _vp), ("caps_base", _vp), ("caps_radius", _vp),
        ("in_collision", _vp), ("sched", _vp),
        ("dp", _vp), ("cnt", _vp), ("split_acc", _vp), ("split_ctr", _vp), ("status", _vp), ("fold_ws", _vp),
        ("fold_stride", C.c_int64), ("host_envs", _vp),
        ("host_level_off", _vp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("substeps", C.c_int32), ("iterations", C.c_int32), ("dt", C.c_float),
        ("gravity", C.c_float * 3), ("damping", C.c_float), ("ground", C.c_float),
        ("dist_mode", C.c_int32), ("dist_k", C.c_float),
        ("vol_mode", C.c_int32), ("vol_k", C.c_float),
        ("nh_on", C.c_int32), ("nh_vol_c", C.c_float), ("nh_dev_c", C.c_float),
        ("drag_ks", C.c_float), ("lambda_legacy", C.c_int32), ("collide", C.c_int32), ("contact_k", C.c_float),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3), ("dirs", _vp), ("height", C.c_int32), ("width", C.c_int32),
        ("mode", C.c_int32), ("light_dir", C.c_float * 3), ("light_intensity", C.c_float),
        ("traversal", C.c_int32),
    ]


class Requests(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("envs", _vp), ("kind", _vp), ("action", _vp), ("x0", _vp), ("tips0", _vp),
        ("lo", _vp), ("hi", _vp), ("base0", _vp),
    ]


class Instrument(C.Structure):
    _fields_ = [
        ("n_env", C.c_int32), ("xform", _vp), ("height", _vp), ("drag_cut_heat", _vp),
        ("drag_cut_heat_update", _vp), ("violation", _vp),
    ]


class BvhJob(C.Structure):
    _fields_ = [
        ("verts", _vp), ("faces", _vp), ("nv", C.c_int32), ("nf", C.c_int32), ("left", _vp), ("first", _vp),
        ("count", _vp), ("level", _vp), ("tri_map", _vp), ("info", _vp), ("scratch", _vp),
    ]


class Parts(C.Structure):
    _fields_ = [
        ("n_parts", C.c_int32), ("max_local", C.c_int32), ("max_solo", C.c_int32), ("reserved", C.c_int32),
        ("parts", _vp), ("env_parts", _vp), ("verts", _vp),
        ("tets16", _vp), ("edges16", _vp), ("tris16", _vp), ("sums", _vp), ("stride", C.c_int64), ("ctr", _vp),
        ("claims", _vp),
        ("n_claims", C.c_int64), ("host_env_parts", _vp),
    ]


REQ_STEP, REQ_RESET = 0, 1

_i32, _f32, _i64 = C.c_int32, C.c_float, C.c_int64
_P = C.POINTER

# name -> (argtypes).  Every function returns int (0 = ok).
SIGNATURES = {
    "ffsrl_frame": [_P(Batch), _P(Params), _vp, _i32, _i32, _vp],
    "ffsrl_frame_envs": [_P(Batch), _P(Params), _vp, _i32, _i32, _i32, _vp],
    "ffsrl_frame_shared": [_P(Batch), _P(Params), _vp, _i32, _i32, _i32, _vp],
    "ffsrl_frame_shared_noalloc": [_P(Batch), _P(Params), _vp, _i32, _i32, _i32, _vp],
    "ffsrl_render_envs": [_P(Batch), _P(Camera), _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ffsrl_apply_requests": [_P(Batch), _P(Requests), _f32, _vp],
    "ffsrl_observe_envs": [_P(Batch), _vp, _i32, _vp, _f32, _vp, _vp],
    "ffsrl_bvh_refit": [_P(Batch), _i32, _i32, _vp],
    "ffsrl_render": [_P(Batch), _P(Camera), _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "ffsrl_vertex_normals": [_P(Batch), _i32, _i32, _vp, _vp],
    "ffsrl_move_tools": [_P(Batch), _vp, _vp, _vp, _vp],
    "ffsrl_grasp_select": [_P(Batch), _f32, _vp, _vp, _vp],
    "ffsrl_predict": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_project_volume": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_project_distance": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_project_neohookean": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_apply": [_P(Batch), _i32, _i32, _vp],
    "ffsrl_grasp": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_collide": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_instrument_actions": [_P(Batch), _P(Instrument), _vp, _f32, _vp],
    "ffsrl_instrument_cartesian": [_P(Batch), _P(Instrument), _vp, _vp, _vp, _vp],
    "ffsrl_instrument_set_rcm": [_P(Batch), _P(Instrument), _vp, _vp, _vp],
    "ffsrl_instrument_update": [_P(Batch), _P(Instrument), _vp],
    "ffsrl_bvh_build_device": [_vp, _i32, _i32, _vp],
    "ffsrl_bounds_integrate": [_P(Batch), _P(Params), _i32, _i32, _vp],
    "ffsrl_zero_lambdas": [_P(Batch), _i32, _i32, _vp],
    "ffsrl_bvh_refit_plain": [_P(Batch), _i32, _i32, _vp],
    "ffsrl_frame_do": [_P(Batch), _P(Params), _i32, _i32, _i32, _vp],
    "ffsrl_frame_parts": [_P(Batch), _P(Params), _P(Parts), _vp, _i32, _i32, _vp],
    "ffsrl_remap": [_vp, _vp, _vp, _i64, _vp],
    "ffsrl_bvh_build": [_vp, _i32, _vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _P(_i32), _P(_i32)],
    "ffsrl_color_constraints": [_vp, _i64, _i32, _i32, _vp, _P(_i32)],
    "ffsrl_order_constraints": [_vp, _i64, _i32, _vp, _i32, _i32, _vp],
    "ffsrl_partition_mesh": [_vp, _i64, _vp, _i64, _vp, _i64, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i64,
                             _P(_i64), _vp, _vp, _vp, _P(_i32)],
    "ffsrl_graph_begin": [_vp],
    "ffsrl_graph_end": [_vp, _P(_vp)],
    "ffsrl_graph_launch": [_vp, _vp],
    "ffsrl_graph_destroy": [_vp],
    "ffsrl_graph_abort": [_vp],
}
# Exported functions that do not return an error code.
PLAIN_SIGNATURES = {
    "ffsrl_abi_version": ([], C.c_int),
    "ffsrl_last_error": ([], C.c_char_p),
    "ffsrl_frame_max_vertices": ([], C.c_int),
    "ffsrl_bvh_scratch_bytes": ([_i32], C.c_int64),
    "ffsrl_xcd_count": ([], C.c_int),          # count > 0, or an error code < 0
    "ffsrl_parts_max_local": ([], C.c_int),
    "ffsrl_parts_max_solo": ([], C.c_int),
}


class FFSRLError(RuntimeError):
    pass


class _Lib:
    def __init__(self):
        self._h = None

    def load(self):
        if self._h is None:
            if not os.path.exists(LIB_PATH):
                raise FFSRLError(
                    f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
            h = C.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = C.c_int
            for name, (args, res) in PLAIN_SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = res
            self._h = h
        return self._h

    def __getattr__(self, name):
        h = self.load()
        fn = getattr(h, name)
        if name in PLAIN_SIGNATURES:
            return fn

        def call(*args):
            rc = fn(*args)
            if rc != 0:
                msg = h.ffsrl_last_error().decode(errors="replace")
                raise FFSRLError(f"{name} failed ({rc}): {msg}")
            return rc

        call.__name__ = name
        return call


lib = _Lib()


def exported_symbols():
    return list(SIGNATURES) + list(PLAIN_SIGNATURES)
