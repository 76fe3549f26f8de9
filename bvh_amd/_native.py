"""ctypes bindings of the native libraries (include/bvh_hip.h, include/bvh_scenes.h).

The HIP library is mandatory: every entry point here raises NativeError when
libbvhhip.so is missing or a call fails.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(HERE, "libbvhhip.so")
LIB_HOST = os.path.join(HERE, "libbvhhost.so")

BVH_OK = 0
STATUS = {
    0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of device memory", 4: "non-finite input",
    5: "no progress", 6: "over memory budget", 7: "no HIP device", 8: "invalid state", 9: "unsupported",
}
ALGO_LBVH, ALGO_LBVH_PLUS, ALGO_PLOC = 0, 1, 2
TRACE_REFERENCE, TRACE_NEAREST = 0, 1
NOALLOC_MAX_TRIANGLES = 262144
BUILD_KEEP_KEYS = 1
BUILD_TUNE = 2

# Byte image of BVHNode (include/bvh_node.h): 48 bytes.
NODE_DTYPE = np.dtype([
    ("bmin", "<f4", (3,)), ("pad0", "<f4"), ("bmax", "<f4", (3,)), ("pad1", "<f4"),
    ("left", "<i4"), ("right", "<i4"), ("pad2", "<i4", (2,)),
])
assert NODE_DTYPE.itemsize == 48


class NativeError(RuntimeError):
    def __init__(self, status, msg=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {msg}" if msg else STATUS.get(status, str(status)))


class BuildParams(C.Structure):
    _fields_ = [("algo", C.c_int32), ("ploc_radius", C.c_int32), ("treelet_rounds", C.c_int32),
                ("radix_bits", C.c_int32), ("flags", C.c_uint32)]


MAX_STAGES = 16


class StageTimes(C.Structure):
    _fields_ = [("count", C.c_int32), ("total_ms", C.c_float), ("ms", C.c_float * MAX_STAGES),
                ("names", (C.c_char * 24) * MAX_STAGES), ("ploc_iterations", C.c_int32), ("reserved", C.c_int32)]

    def as_dict(self):
        return {self.names[i].value.decode(): float(self.ms[i]) for i in range(self.count)}


class Footprint(C.Structure):
    _fields_ = [("mesh", C.c_uint64), ("tree", C.c_uint64), ("build_workspace", C.c_uint64),
                ("trace_workspace", C.c_uint64)]


class BatchInfo(C.Structure):
    _fields_ = [("meshes", C.c_uint32), ("triangles", C.c_uint32), ("max_triangles", C.c_uint32),
                ("algo", C.c_int32), ("radix_bits", C.c_int32), ("sort_passes", C.c_int32), ("key_bits", C.c_int32),
                ("ploc_captured_iterations", C.c_int32), ("ploc_iterations", C.c_int32),
                ("ploc_nn_grid", C.c_int32), ("ploc_merge_grid", C.c_int32), ("graph_nodes", C.c_int32),
                ("runs", C.c_int32), ("eager_continuations", C.c_int32), ("sort_block", C.c_int32),
                ("front_chunk", C.c_int32), ("tune_count", C.c_int32),
                ("tune_ms", C.c_float * 16), ("tune_radix_bits", C.c_int32 * 16), ("tune_nn_grid", C.c_int32 * 16),
                ("tune_sort_block", C.c_int32 * 16), ("tune_front_chunk", C.c_int32 * 16),
                ("tune_merge_grid", C.c_int32 * 16)]
    _TUNE = ("radix_bits", "nn_grid", "sort_block", "front_chunk", "merge_grid")

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_ if not name.startswith("tune_")}
        d["tune_count"] = int(self.tune_count)
        d["tune"] = [dict({"ms": float(self.tune_ms[i])},
                          **{k: int(getattr(self, "tune_" + k)[i]) for k in self._TUNE})
                     for i in range(min(self.tune_count, 16))]
        return d


_hip = None
_host = None

_vp = C.c_void_p
_u32 = C.c_uint32
_i32 = C.c_int32
_u64 = C.c_uint64
_f = C.c_float
_sz = C.c_size_t
_pp = C.POINTER(C.c_void_p)

_HIP_SIGS = {
    "bvh_device_count": (_i32, [C.POINTER(C.c_int)]),
    "bvh_ctx_create": (_i32, [C.c_int, _pp]),
    "bvh_ctx_destroy": (_i32, [_vp]),
    "bvh_last_error": (C.c_char_p, [_vp]),
    "bvh_status_string": (C.c_char_p, [C.c_int]),
    "bvh_device_memory": (_i32, [_vp, C.POINTER(_u64), C.POINTER(_u64)]),
    "bvh_device_alloc": (_i32, [_vp, _sz, _pp]),
    "bvh_device_free": (_i32, [_vp, _vp]),
    "bvh_copy_h2d": (_i32, [_vp, _vp, _vp, _sz]),
    "bvh_copy_d2h": (_i32, [_vp, _vp, _vp, _sz]),
    "bvh_host_alloc_pinned": (_i32, [_sz, _pp]),
    "bvh_host_free_pinned": (_i32, [_vp]),
    "bvh_mesh_upload": (_i32, [_vp, C.POINTER(_vp), _u32, _pp]),
    "bvh_mesh_free": (_i32, [_vp, _vp]),
    "bvh_mesh_release": (_i32, [_vp, _vp]),
    "bvh_device_cache_info": (_i32, [_i32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "bvh_mesh_bytes": (_i32, [_vp, C.POINTER(C.c_uint64)]),
    "bvh_tree_bytes": (_i32, [_vp, C.POINTER(C.c_uint64)]),
    "bvh_device_cache_trim": (_i32, [_i32]),
    "bvh_device_cache_set_limit": (_i32, [_i32, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bvh_mesh_size": (_i32, [_vp, C.POINTER(_u32)]),
    "bvh_mesh_check": (_i32, [_vp]),
    "bvh_tree_create": (_i32, [_vp, _pp]),
    "bvh_tree_free": (_i32, [_vp, _vp]),
    "bvh_tree_release": (_i32, [_vp, _vp]),
    "bvh_tree_info": (_i32, [_vp, C.POINTER(_u32), C.POINTER(_u32)]),
    "bvh_tree_download": (_i32, [_vp, _vp, _vp, _vp]),
    "bvh_tree_from_nodes": (_i32, [_vp, _vp, _u32, _vp]),
    "bvh_tree_set_indices": (_i32, [_vp, _vp, _vp, _u32]),
    "bvh_build": (_i32, [_vp, _vp, C.POINTER(BuildParams), _vp, C.POINTER(StageTimes)]),
    "bvh_build_lbvh_noalloc": (_i32, [_vp, C.POINTER(_vp), _u32, _vp, _vp]),
    "bvh_footprint_estimate": (_i32, [_i32, _u32, _u32, _u32, C.POINTER(Footprint)]),
    "bvh_ctx_held_bytes": (_i32, [_vp, C.POINTER(_u64), C.POINTER(_u64)]),
    "bvh_debug_morton": (_i32, [_vp, _vp, _i32, _vp]),
    "bvh_debug_sorted_keys": (_i32, [_vp, _vp, _vp, _u32]),
    "bvh_sort_pairs_u64": (_i32, [_vp, _vp, _vp, _u32, C.c_int, C.c_int, C.c_int, C.POINTER(_f)]),
    "bvh_debug_stream_copy": (_i32, [_vp, _vp, _vp, _sz, C.POINTER(_f)]),
    "bvh_build_batched": (_i32, [_vp, C.POINTER(_vp), C.POINTER(_vp), _u32, C.POINTER(BuildParams),
                                 C.POINTER(StageTimes)]),
    "bvh_batch_create": (_i32, [_vp, C.POINTER(_vp), C.POINTER(_vp), _u32, C.POINTER(BuildParams), _pp]),
    "bvh_batch_run": (_i32, [_vp, C.POINTER(_f)]),
    "bvh_batch_get_info": (_i32, [_vp, C.POINTER(BatchInfo)]),
    "bvh_batch_free": (_i32, [_vp]),
    "bvh_tree_prepare": (_i32, [_vp, _vp]),
    "bvh_ctx_set_trace_order": (_i32, [_vp, _i32]),
    "bvh_trace": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, C.POINTER(_f)]),
    "bvh_trace_host": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, C.POINTER(_f)]),
    "bvh_build_async": (_i32, [_vp, _vp, C.POINTER(BuildParams), _vp]),
    "bvh_trace_async": (_i32, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    "bvh_sync": (_i32, [_vp]),
}

_HOST_SIGS = {
    "bvh_rng_u64": (_i32, [_u64, _u32, _vp]),
    "bvh_scene_uniform": (_i32, [_u64, _u32, C.POINTER(_vp)]),
    "bvh_scene_blobs": (_i32, [_u64, _u32, C.POINTER(_vp)]),
    "bvh_scene_strands": (_i32, [_u64, _u32, _u32, C.POINTER(_vp)]),
    "bvh_scene_batch_counts": (_i32, [_u64, _u32, _u32, _u32, _vp]),
    "bvh_scene_batch_mesh": (_i32, [_u64, _u32, _u32, C.POINTER(_vp)]),
    "bvh_rays_camera": (_i32, [_vp, _vp, _vp, _f, _u32, _u32, _vp]),
    "bvh_rays_sphere_camera": (_i32, [_vp, _f, _u32, _u32, _vp]),
    "bvh_rays_sphere_to_box": (_i32, [_u64, _vp, _u32, _vp]),
    "bvh_rays_uniform_cube": (_i32, [_u64, _u32, _vp]),
    "bvh_fnv1a64": (_u64, [_vp, _sz, _u64]),
    "bvh_tree_file_write": (_i32, [C.c_char_p, _vp, _u32, _vp, _u32]),
    "bvh_tree_validate": (_i32, [_vp, _u32, _vp, _u32]),
    "bvh_tree_file_read": (_i32, [C.c_char_p, _vp, C.POINTER(_u32), _vp, C.POINTER(_u32)]),
    "bvh_tree_save": (_i32, [_vp, _vp, C.c_char_p]),
    "bvh_tree_load": (_i32, [_vp, C.c_char_p, _vp]),
    "bvh_tree_export_viz": (_i32, [_vp, _vp, C.c_char_p]),
    "bvh_nodes_export_obj": (_i32, [_vp, _u32, C.c_char_p, _i32]),
    "bvh_tree_export_obj": (_i32, [_vp, _vp, C.c_char_p, _i32]),
    "bvh_obj_parse": (_i32, [C.c_char_p, _i32, _pp]),
    "bvh_obj_info": (_i32, [_vp, C.POINTER(_u32), C.POINTER(_u64)]),
    "bvh_obj_fill": (_i32, [_vp, C.POINTER(_vp)]),
    "bvh_obj_free": (_i32, [_vp]),
    "bvh_obj_last_error": (C.c_char_p, []),
    "bvh_mesh_upload_obj": (_i32, [_vp, C.c_char_p, _i32, _pp]),
    "bvh_render_ppm": (_i32, [_vp, _u32, C.POINTER(_vp), _u32, _i32, _i32, _vp, _i32, _f, C.c_char_p, _vp]),
}


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def hip():
    """The gfx950 kernel library; raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(LIB_HIP):
            raise NativeError(2, f"{LIB_HIP} is missing: run __graft_entry__.build() (python -m bvh_amd.build)")
        lib = C.CDLL(LIB_HIP, mode=C.RTLD_GLOBAL)
        _bind(lib, _HIP_SIGS)
        _hip = lib
    return _hip


def host():
    global _host
    if _host is None:
        hip()
        if not os.path.exists(LIB_HOST):
            raise NativeError(2, f"{LIB_HOST} is missing: run python -m bvh_amd.build")
        lib = C.CDLL(LIB_HOST, mode=C.RTLD_GLOBAL)
        _bind(lib, _HOST_SIGS)
        _host = lib
    return _host


def ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def soa_ptrs(soa: np.ndarray):
    """soa: (9, n) float32 C-contiguous -> array of 9 pointers."""
    arr = (C.c_void_p * 9)()
    for k in range(9):
        arr[k] = soa[k].ctypes.data
    return arr


def device_count() -> int:
    c = C.c_int(0)
    hip().bvh_device_count(C.byref(c))
    return c.value
