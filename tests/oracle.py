"""ctypes view of oracle/liboracle.so (the CPU restatement of the reference).

Test infrastructure: only tests, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.environ.get("BVH_ORACLE_SO") or os.path.join(ROOT, "oracle", "liboracle.so")

NODE_DTYPE = np.dtype([
    ("bmin", "<f4", (3,)), ("pad0", "<f4"), ("bmax", "<f4", (3,)), ("pad1", "<f4"),
    ("left", "<i4"), ("right", "<i4"), ("pad2", "<i4", (2,)),
])

_lib = None
_vp = C.c_void_p
_i64 = C.c_int64


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True)
        L = C.CDLL(ORACLE_SO)
        sigs = {
            "orc_tri_boxes": (None, [_vp, _i64, _vp]),
            "orc_scene_box": (None, [_vp, _i64, _vp]),
            "orc_expand10": (C.c_uint32, [C.c_uint32]),
            "orc_expand21": (C.c_uint64, [C.c_uint64]),
            "orc_morton30_point": (C.c_uint32, [C.c_float, C.c_float, C.c_float]),
            "orc_morton60_point": (C.c_uint64, [C.c_float, C.c_float, C.c_float]),
            "orc_morton30": (None, [_vp, _i64, _vp]),
            "orc_morton60": (None, [_vp, _i64, _vp]),
            "orc_karras30": (None, [_vp, _vp, _i64, _vp, _vp, _vp]),
            "orc_build_lbvh": (None, [_vp, _i64, _vp, _vp, _vp]),
            "orc_build_lbvh_plus": (None, [_vp, _i64, C.c_int, _vp, _vp, _vp]),
            "orc_optimize_treelet": (None, [_vp, _vp, _i64, C.c_int]),
            "orc_build_ploc": (C.c_int, [_vp, _i64, C.c_int, _vp, _vp, _vp, C.c_int32, _vp]),
            "orc_trace": (None, [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp]),
            "orc_trace_omp": (None, [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp]),
            "orc_trace_nearest": (None, [_vp, _i64, _vp, _vp, _i64, _vp, _vp, _vp]),
            "orc_trace_brute": (None, [_vp, _i64, _vp, _i64, _vp, _vp]),
            "orc_sah": (C.c_float, [_vp, _i64, C.c_float, C.c_float]),
            "orc_depth_stats": (None, [_vp, _i64, _vp, _vp, _vp]),
            "orc_omp_threads": (C.c_int, []),
            "orc_build_lbvh_omp": (None, [_vp, _i64, _vp, _vp, _vp]),
            "orc_build_lbvh_plus_omp": (None, [_vp, _i64, C.c_int, _vp, _vp]),
            "orc_build_ploc_omp": (C.c_int, [_vp, _i64, C.c_int, _vp, _vp, _vp, C.c_int32, _vp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _soa(soa):
    arr = (C.c_void_p * 9)()
    for k in range(9):
        arr[k] = soa[k].ctypes.data
    return arr


def _as_soa(mesh_or_soa):
    soa = getattr(mesh_or_soa, "soa", mesh_or_soa)
    return np.ascontiguousarray(soa, dtype=np.float32)


def tri_boxes(m):
    soa = _as_soa(m)
    n = soa.shape[1]
    out = np.zeros((n, 8), dtype=np.float32)
    lib().orc_tri_boxes(_soa(soa), n, out.ctypes.data)
    return out


def morton30(m):
    soa = _as_soa(m)
    out = np.zeros(soa.shape[1], dtype=np.uint32)
    lib().orc_morton30(_soa(soa), soa.shape[1], out.ctypes.data)
    return out


def morton60(m):
    soa = _as_soa(m)
    out = np.zeros(soa.shape[1], dtype=np.uint64)
    lib().orc_morton60(_soa(soa), soa.shape[1], out.ctypes.data)
    return out


def karras30(codes, idx):
    codes = np.ascontiguousarray(codes, np.uint32)
    idx = np.ascontiguousarray(idx, np.uint32)
    n = codes.size
    left = np.zeros(max(n - 1, 1), np.uint32)
    right = np.zeros(max(n - 1, 1), np.uint32)
    parent = np.zeros(2 * n - 1, np.uint32)
    lib().orc_karras30(codes.ctypes.data, idx.ctypes.data, n, left.ctypes.data, right.ctypes.data, parent.ctypes.data)
    return left[: n - 1], right[: n - 1], parent


def build_lbvh(m):
    soa = _as_soa(m)
    n = soa.shape[1]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    idx = np.zeros(n, dtype=np.uint32)
    keys = np.zeros(n, dtype=np.uint64)
    lib().orc_build_lbvh(_soa(soa), n, nodes.ctypes.data, idx.ctypes.data, keys.ctypes.data)
    return nodes, idx, keys


def build_lbvh_omp(m):
    soa = _as_soa(m)
    n = soa.shape[1]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    idx = np.zeros(n, dtype=np.uint32)
    ms = np.zeros(5, dtype=np.float64)
    lib().orc_build_lbvh_omp(_soa(soa), n, nodes.ctypes.data, idx.ctypes.data, ms.ctypes.data)
    return nodes, idx, ms


def build_lbvh_plus(m, rounds=1):
    soa = _as_soa(m)
    n = soa.shape[1]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    idx = np.zeros(n, dtype=np.uint32)
    codes = np.zeros(n, dtype=np.uint64)
    lib().orc_build_lbvh_plus(_soa(soa), n, rounds, nodes.ctypes.data, idx.ctypes.data, codes.ctypes.data)
    return nodes, idx, codes


def build_lbvh_plus_omp(m, rounds=1):
    """OpenMP LBVH+ (treelet rounds on threads); equal to build_lbvh_plus."""
    soa = _as_soa(m)
    n = soa.shape[1]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    idx = np.zeros(n, dtype=np.uint32)
    lib().orc_build_lbvh_plus_omp(_soa(soa), n, rounds, nodes.ctypes.data, idx.ctypes.data)
    return nodes, idx


def build_ploc_omp(m, radius=25, max_iters=4096):
    """OpenMP PLOC; equal to build_ploc (status, nodes, indices, clusters
    entering each iteration)."""
    soa = _as_soa(m)
    n = soa.shape[1]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    idx = np.zeros(n, dtype=np.uint32)
    iters = np.zeros(max_iters, dtype=np.int32)
    nit = C.c_int32(0)
    st = lib().orc_build_ploc_omp(_soa(soa), n, radius, nodes.ctypes.data, idx.ctypes.data, iters.ctypes.data,
                                  max_iters, C.byref(nit))
    return st, nodes, idx, iters[: min(nit.value, max_iters)]


def optimize_treelet(nodes, parents, root):
    nodes = np.ascontiguousarray(nodes.copy(), dtype=NODE_DTYPE)
    parents = np.ascontiguousarray(parents.copy(), dtype=np.uint32)
    lib().orc_optimize_treelet(nodes.ctypes.data, parents.ctypes.data, nodes.size, int(root))
    return nodes, parents


def build_ploc(m, radius=25, max_iters=4096):
    soa = _as_soa(m)
    n = soa.shape[1]
    nodes = np.zeros(2 * n - 1, dtype=NODE_DTYPE)
    idx = np.zeros(n, dtype=np.uint32)
    iters = np.zeros(max_iters, dtype=np.int32)
    nit = C.c_int32(0)
    st = lib().orc_build_ploc(_soa(soa), n, radius, nodes.ctypes.data, idx.ctypes.data, iters.ctypes.data,
                              max_iters, C.byref(nit))
    return st, nodes, idx, iters[: min(nit.value, max_iters)]


def trace(nodes, m, rays, omp=False, nearest=False):
    """Closest hit per ray (t, prim, counters): the reference order, or the
    nearest-first order (least (t, index) hit) when nearest."""
    soa = _as_soa(m)
    rays = np.ascontiguousarray(rays, np.float32)
    nr = rays.shape[0]
    t = np.zeros(nr, np.float32)
    prim = np.zeros(nr, np.int32)
    cnt = np.zeros((nr, 3), np.int32)
    f = lib().orc_trace_nearest if nearest else (lib().orc_trace_omp if omp else lib().orc_trace)
    f(nodes.ctypes.data, nodes.size, _soa(soa), rays.ctypes.data, nr, t.ctypes.data, prim.ctypes.data,
      cnt.ctypes.data)
    return t, prim, cnt


def trace_brute(m, rays):
    soa = _as_soa(m)
    rays = np.ascontiguousarray(rays, np.float32)
    nr = rays.shape[0]
    t = np.zeros(nr, np.float32)
    prim = np.zeros(nr, np.int32)
    lib().orc_trace_brute(_soa(soa), soa.shape[1], rays.ctypes.data, nr, t.ctypes.data, prim.ctypes.data)
    return t, prim


def sah(nodes, ct=3.0, ci=2.0):
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    return float(lib().orc_sah(nodes.ctypes.data, nodes.size, ct, ci))


def depth_stats(nodes):
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    md = C.c_int32(0)
    avg = C.c_double(0)
    lv = C.c_int64(0)
    lib().orc_depth_stats(nodes.ctypes.data, nodes.size, C.byref(md), C.byref(avg), C.byref(lv))
    return md.value, avg.value, lv.value


def nodes_equal(a, b):
    """Compares the meaningful fields of two node arrays (padding ignored); returns first mismatch or None."""
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    for f in ("left", "right"):
        bad = np.nonzero(a[f] != b[f])[0]
        if bad.size:
            return f"{f} differs at node {bad[0]}: {a[f][bad[0]]} vs {b[f][bad[0]]}"
    for f in ("bmin", "bmax"):
        av = a[f].view(np.uint32)
        bv = b[f].view(np.uint32)
        bad = np.nonzero((av != bv).any(axis=1))[0]
        if bad.size:
            return f"{f} differs at node {bad[0]}: {a[f][bad[0]]} vs {b[f][bad[0]]}"
    return None
