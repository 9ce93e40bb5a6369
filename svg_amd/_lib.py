"""ctypes binding of libsvg_amd.so (the C ABI declared in include/svg_amd.h).

The product path always runs the HIP library; if it is not built this module
raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SVG_AMD_LIB names another build of the same library (a sanitizer build
# of the host code, for instance); by default the in-tree build is loaded.
LIB_PATH = os.environ.get("SVG_AMD_LIB") or os.path.join(HERE, "libsvg_amd.so")

SVG_MODE_PLAIN = 0
SVG_MODE_BATCHED = 1


class SvgError(RuntimeError):
    pass


class BuildParams(C.Structure):
    _fields_ = [("K", C.c_int), ("split_coef", C.c_uint32), ("vert_coef", C.c_uint32),
                ("mode", C.c_int), ("plain_blocks", C.c_int), ("plain_threads", C.c_int),
                ("lanes", C.c_int), ("prop_block", C.c_int), ("chunk", C.c_int),
                ("vert_cap", C.c_int), ("split_cap", C.c_int), ("autotune", C.c_int),
                ("lanes_per_wave", C.c_int), ("queue_cap", C.c_int), ("source_order", C.c_int)]


class BuildStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("prop_seconds", C.c_double),
                ("extract_seconds", C.c_double), ("redo_seconds", C.c_double),
                ("alloc_seconds", C.c_double), ("arcs", C.c_longlong), ("redone", C.c_int),
                ("flagged_drop", C.c_int), ("flagged_overflow", C.c_int),
                ("flagged_error", C.c_int), ("lanes", C.c_int), ("prop_block", C.c_int),
                ("chunk", C.c_int), ("vert_cap", C.c_int), ("split_cap", C.c_int),
                ("tuned", C.c_int), ("counters", C.c_longlong * 16), ("lanes_per_wave", C.c_int),
                ("queue_cap", C.c_int), ("pseudo_queue_cap", C.c_int), ("fallback_plain", C.c_int),
                ("source_order", C.c_int)]


class ServerCfg(C.Structure):
    _fields_ = [("max_inflight", C.c_int), ("mem_budget", C.c_longlong), ("max_batch", C.c_int),
                ("delta_scale", C.c_double), ("rounds_per_sync", C.c_int),
                ("msad_delta_scale", C.c_double), ("streams", C.c_int), ("chain_rows", C.c_int),
                ("queue_cap", C.c_longlong)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("s", C.c_int), ("t", C.c_int),
                ("src", C.POINTER(C.c_int32)), ("w", C.POINTER(C.c_double)), ("ns", C.c_int),
                ("dist_out", C.POINTER(C.c_double)), ("srcid_out", C.POINTER(C.c_int32))]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_longlong), ("refused", C.c_longlong), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("qps", C.c_double), ("max_batch_seen", C.c_int), ("launches", C.c_longlong),
                ("n_kind", C.c_longlong * 2), ("rounds_kind", C.c_longlong * 2),
                ("settled_kind", C.c_longlong * 2), ("relaxed_kind", C.c_longlong * 2),
                ("slots", C.c_int), ("slot_bytes", C.c_longlong)]


COUNTER_NAMES = ["win_push", "win_pop", "pseudo_push", "pseudo_pop", "propagate",
                 "child_built", "split_write", "vert_write", "fan_halfedge", "pseudo_neigh",
                 "win_drop", "pseudo_drop", "bt_hops", "reached", "arcs", "errors"]

# Symbols the C ABI exports (include/svg_amd.h), used by the load check.
EXPORTS = [
    "svg_init", "svg_last_error", "svg_device_name", "svg_mesh_from_arrays", "svg_mesh_load_obj",
    "svg_trace_arc", "svg_geodesic_path",
    "svg_mesh_upload", "svg_mesh_info", "svg_mesh_export", "svg_mesh_free",
    "svg_build_params_default", "svg_build", "svg_build_stats_get", "svg_build_status",
    "svg_graph_info", "svg_graph_to_host", "svg_graph_save", "svg_graph_load", "svg_graph_free",
    "svg_sssd", "svg_msad", "svg_msad_noalloc", "svg_sssd_point", "svg_sssd_point_batch", "svg_sssd_batch", "svg_msad_batch",
    "svg_server_start", "svg_server_submit", "svg_server_submit_cb",
    "svg_server_wait", "svg_server_stats_get", "svg_server_stop", "svg_server_replay",
]

# svg_done_fn: void (*)(long long ticket, int status, double latency_ms, void *user)
DONE_FN = C.CFUNCTYPE(None, C.c_longlong, C.c_int, C.c_double, C.c_void_p)

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_bp = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p


def lib():
    """Load libsvg_amd.so (built by __graft_entry__.build / svg_amd.build_lib)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SvgError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    L.svg_last_error.restype = C.c_char_p
    L.svg_init.argtypes = [C.c_int]
    L.svg_device_name.argtypes = [C.c_char_p, C.c_int]
    L.svg_mesh_from_arrays.argtypes = [_dp, C.c_int, _ip, C.c_int, C.POINTER(_vp)]
    L.svg_mesh_load_obj.argtypes = [C.c_char_p, C.POINTER(_vp)]
    L.svg_mesh_upload.argtypes = [_vp]
    L.svg_mesh_info.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.svg_mesh_export.argtypes = [_vp, _ip, _dp, _dp, _ip, _bp]
    L.svg_mesh_free.argtypes = [_vp]
    L.svg_build_params_default.argtypes = [C.POINTER(BuildParams), C.c_int]
    L.svg_build_params_default.restype = None
    L.svg_build.argtypes = [_vp, C.POINTER(BuildParams), C.POINTER(_vp)]
    L.svg_build_stats_get.argtypes = [_vp, C.POINTER(BuildStats)]
    L.svg_build_status.argtypes = [_vp, _up]
    L.svg_graph_info.argtypes = [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_longlong)]
    L.svg_graph_to_host.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.svg_graph_save.argtypes = [_vp, C.c_char_p]
    L.svg_graph_load.argtypes = [_vp, C.c_int, C.c_char_p, C.POINTER(_vp)]
    L.svg_graph_free.argtypes = [_vp]
    L.svg_sssd.argtypes = [_vp, C.c_int, C.c_int, C.POINTER(C.c_double), _vp, C.c_int,
                           C.POINTER(C.c_int)]
    L.svg_msad.argtypes = [_vp, _ip, _dp, C.c_int, _dp, _vp, _vp]
    L.svg_msad_noalloc.argtypes = [_vp, _ip, _dp, C.c_int, _dp, _vp, _vp]
    L.svg_sssd_batch.argtypes = [_vp, _ip, C.c_int, _dp]
    L.svg_msad_batch.argtypes = [_vp, _ip, _dp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"),
                                 C.c_int, _dp, _vp]
    L.svg_sssd_point_batch.argtypes = [_vp, C.c_int, _ip, _dp, _dp, _ip, _dp, _ip, _ip]
    L.svg_sssd_point.argtypes = [_vp, C.c_int, _dp, C.c_int, _dp, C.POINTER(C.c_double), _ip, _dp,
                                 C.POINTER(C.c_int32), C.POINTER(C.c_int), _vp, C.c_int, C.POINTER(C.c_int)]
    L.svg_trace_arc.argtypes = [_vp, C.c_int, C.c_int, C.c_int32, C.c_double, C.c_int32, _dp, C.c_int,
                                C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.svg_geodesic_path.argtypes = [_vp, _ip, C.c_int, _dp, C.c_int, C.POINTER(C.c_int)]
    if hasattr(L, "svg_server_start"):
        L.svg_server_start.argtypes = [_vp, C.POINTER(ServerCfg), C.POINTER(_vp)]
        L.svg_server_submit.argtypes = [_vp, C.POINTER(Request)]
        L.svg_server_submit.restype = C.c_longlong
        L.svg_server_submit_cb.argtypes = [_vp, C.POINTER(Request), DONE_FN, C.c_void_p]
        L.svg_server_submit_cb.restype = C.c_longlong
        L.svg_server_wait.argtypes = [_vp, C.c_longlong, C.POINTER(C.c_double)]
        L.svg_server_stats_get.argtypes = [_vp, C.POINTER(ServerStats)]
        L.svg_server_stop.argtypes = [_vp]
        L.svg_server_replay.argtypes = [_vp, C.POINTER(Request), _dp, C.c_int, _dp]
    _lib = L
    return L


def check(rc: int):
    if rc < 0:
        raise SvgError(f"svg error {rc}: {lib().svg_last_error().decode(errors='replace')}")
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
