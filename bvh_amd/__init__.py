"""bvh_amd — MI355X-native (gfx950) BVH build and trace engine.

Python mirror of the reference builder interface (include/bvh_builder.h of the
reference: getName/build/getNodes/getIndices/getLastBuildTimeMS/
getTimingBreakdown) over the C ABI of include/bvh_hip.h.  The kernels run on
the GPU through libbvhhip.so; there is no CPU path in this package.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import NODE_DTYPE, NativeError, ALGO_LBVH, ALGO_LBVH_PLUS, ALGO_PLOC  # noqa: F401
from ._native import TRACE_REFERENCE, TRACE_NEAREST  # noqa: F401

__all__ = [
    "Mesh", "Context", "DeviceMesh", "DeviceBuffer", "Tree", "LBVHBuilder", "LBVHPlusBuilder", "PLOCBuilder", "NODE_DTYPE",
    "NativeError", "TRACE_REFERENCE", "TRACE_NEAREST", "rays_sphere_camera", "rays_sphere_to_box", "rays_uniform_cube", "rays_camera", "Batch",
]


class Mesh:
    """Host triangle soup in SoA form: soa[k] for k = v0x v0y v0z v1x v1y v1z v2x v2y v2z."""

    def __init__(self, soa: np.ndarray):
        soa = np.ascontiguousarray(soa, dtype=np.float32)
        if soa.ndim != 2 or soa.shape[0] != 9:
            raise ValueError("soa must have shape (9, n)")
        self.soa = soa

    def __len__(self):
        return self.soa.shape[1]

    @property
    def n(self):
        return self.soa.shape[1]

    @staticmethod
    def _empty(n):
        return np.zeros((9, n), dtype=np.float32)

    @classmethod
    def uniform(cls, seed: int, n: int) -> "Mesh":
        soa = cls._empty(n)
        N.host().bvh_scene_uniform(seed, n, N.soa_ptrs(soa))
        return cls(soa)

    @classmethod
    def blobs(cls, seed: int, n: int) -> "Mesh":
        soa = cls._empty(n)
        N.host().bvh_scene_blobs(seed, n, N.soa_ptrs(soa))
        return cls(soa)

    @classmethod
    def strands(cls, seed: int, strands: int, steps: int) -> "Mesh":
        soa = cls._empty(strands * steps)
        N.host().bvh_scene_strands(seed, strands, steps, N.soa_ptrs(soa))
        return cls(soa)

    @classmethod
    def batch_mesh(cls, seed: int, index: int, n: int) -> "Mesh":
        soa = cls._empty(n)
        N.host().bvh_scene_batch_mesh(seed, index, n, N.soa_ptrs(soa))
        return cls(soa)

    @classmethod
    def from_obj(cls, path: str, threads: int = 0) -> "Mesh":
        """Parses an OBJ file with the reference loader's rules (include/bvh_obj.h)."""
        h = C.c_void_p()
        st = N.host().bvh_obj_parse(path.encode(), threads, C.byref(h))
        if st != 0:
            raise NativeError(st, N.host().bvh_obj_last_error().decode())
        try:
            nt, nv = C.c_uint32(), C.c_uint64()
            N.host().bvh_obj_info(h, C.byref(nt), C.byref(nv))
            soa = cls._empty(nt.value)
            if nt.value:
                N.host().bvh_obj_fill(h, N.soa_ptrs(soa))
            return cls(soa)
        finally:
            N.host().bvh_obj_free(h)

    def checksum(self) -> int:
        h = 14695981039346656037
        for k in range(9):
            h = N.host().bvh_fnv1a64(self.soa[k].ctypes.data, self.soa[k].nbytes, h)
        return h

    def bounds(self) -> np.ndarray:
        v = self.soa.reshape(3, 3, -1)
        return np.concatenate([v.min(axis=(0, 2)), v.max(axis=(0, 2))]).astype(np.float32)


def tree_file_write(path: str, nodes: np.ndarray, indices: np.ndarray) -> None:
    """Host-only writer of the tree file format (no device)."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    indices = np.ascontiguousarray(indices, dtype=np.uint32)
    st = N.host().bvh_tree_file_write(path.encode(), nodes.ctypes.data, nodes.size, indices.ctypes.data,
                                      indices.size)
    if st != 0:
        raise NativeError(st, f"tree file write {path}")


def tree_file_read(path: str):
    nn, np_ = C.c_uint32(), C.c_uint32()
    st = N.host().bvh_tree_file_read(path.encode(), None, C.byref(nn), None, C.byref(np_))
    if st != 0:
        raise NativeError(st, f"tree file read {path}")
    nodes = np.zeros(nn.value, dtype=NODE_DTYPE)
    idx = np.zeros(np_.value, dtype=np.uint32)
    st = N.host().bvh_tree_file_read(path.encode(), nodes.ctypes.data, C.byref(nn), idx.ctypes.data, C.byref(np_))
    if st != 0:
        raise NativeError(st, f"tree file read {path}")
    return nodes, idx


def render_ppm(nodes: np.ndarray, mesh: Mesh, width: int, height: int, eye, look_at, up=(0, 1, 0), fov=60.0,
               mode: int = 0, fixed_heatmap_max: float = 0.0, path: str | None = None) -> dict:
    """renderImage of the reference API (include/render_engine.h) through its
    C entry point: traces on the GPU, shades on the host, writes a PPM."""
    nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
    cam = np.array([*eye, *look_at, *up, fov], np.float32)
    stats = np.zeros(5, np.float32)
    st = N.host().bvh_render_ppm(nodes.ctypes.data, nodes.size, N.soa_ptrs(mesh.soa), mesh.n, width, height,
                                 cam.ctypes.data, mode, fixed_heatmap_max, path.encode() if path else None,
                                 stats.ctypes.data)
    if st != 0:
        raise NativeError(st, "render")
    return {"render_ms": float(stats[0]), "avg_nodes_visited": float(stats[1]), "avg_aabb_tests": float(stats[2]),
            "avg_tri_tests": float(stats[3]), "max_nodes_visited": float(stats[4])}


def batch_counts(seed: int, meshes: int, lo: int, hi: int) -> np.ndarray:
    out = np.zeros(meshes, dtype=np.uint32)
    N.host().bvh_scene_batch_counts(seed, meshes, lo, hi, out.ctypes.data)
    return out


def rays_camera(eye, look_at, up, fov, width, height) -> np.ndarray:
    out = np.zeros((width * height, 6), dtype=np.float32)
    e = np.asarray(eye, np.float32); la = np.asarray(look_at, np.float32); u = np.asarray(up, np.float32)
    N.host().bvh_rays_camera(e.ctypes.data, la.ctypes.data, u.ctypes.data, fov, width, height, out.ctypes.data)
    return out


def rays_sphere_camera(box, width=1024, height=768, fov=60.0) -> np.ndarray:
    out = np.zeros((width * height, 6), dtype=np.float32)
    b = np.asarray(box, np.float32)
    N.host().bvh_rays_sphere_camera(b.ctypes.data, fov, width, height, out.ctypes.data)
    return out


def rays_sphere_to_box(seed, box, n) -> np.ndarray:
    out = np.zeros((n, 6), dtype=np.float32)
    b = np.asarray(box, np.float32)
    N.host().bvh_rays_sphere_to_box(seed, b.ctypes.data, n, out.ctypes.data)
    return out


def rays_uniform_cube(seed, n) -> np.ndarray:
    out = np.zeros((n, 6), dtype=np.float32)
    N.host().bvh_rays_uniform_cube(seed, n, out.ctypes.data)
    return out


class Context:
    """One HIP stream + pooled workspace on one device (bvh_ctx_t)."""

    def __init__(self, device: int = 0):
        self._lib = N.hip()
        h = C.c_void_p()
        st = self._lib.bvh_ctx_create(device, C.byref(h))
        if st != N.BVH_OK:
            raise NativeError(st, "bvh_ctx_create")
        self.h = h
        self.device = device

    def _check(self, st, what=""):
        if st != N.BVH_OK:
            raise NativeError(st, f"{what}: {self._lib.bvh_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self._lib.bvh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, mesh: Mesh) -> "DeviceMesh":
        h = C.c_void_p()
        self._check(self._lib.bvh_mesh_upload(self.h, N.soa_ptrs(mesh.soa), mesh.n, C.byref(h)), "mesh upload")
        return DeviceMesh(self, h, mesh.n)

    def upload_obj(self, path: str, threads: int = 0) -> "DeviceMesh":
        """Parses an OBJ file into pinned staging memory and uploads it (bvh_mesh_upload_obj)."""
        h = C.c_void_p()
        st = N.host().bvh_mesh_upload_obj(self.h, path.encode(), threads, C.byref(h))
        if st != N.BVH_OK:
            raise NativeError(st, N.host().bvh_obj_last_error().decode() or self._lib.bvh_last_error(self.h).decode())
        n = C.c_uint32()
        self._lib.bvh_mesh_size(h, C.byref(n))
        return DeviceMesh(self, h, n.value)

    def tree(self) -> "Tree":
        h = C.c_void_p()
        self._check(self._lib.bvh_tree_create(self.h, C.byref(h)), "tree create")
        return Tree(self, h)

    def build(self, dmesh: "DeviceMesh", tree: "Tree", algo=ALGO_LBVH, radius=25, rounds=1, radix_bits=0,
              keep_keys=False, timing=True) -> dict:
        """Builds `tree` from `dmesh`.  timing=False records no per-stage events
        (the returned times are then empty)."""
        p = N.BuildParams(algo, radius, rounds, radix_bits, N.BUILD_KEEP_KEYS if keep_keys else 0)
        if not timing:
            self._check(self._lib.bvh_build(self.h, dmesh.h, C.byref(p), tree.h, None), "build")
            return {"total_ms": None, "stages": {}, "ploc_iterations": 0}
        t = N.StageTimes()
        self._check(self._lib.bvh_build(self.h, dmesh.h, C.byref(p), tree.h, C.byref(t)), "build")
        return {"total_ms": float(t.total_ms), "stages": t.as_dict(), "ploc_iterations": int(t.ploc_iterations)}

    def batch(self, dmeshes, trees, algo=ALGO_LBVH, radius=25, radix_bits=0, tune=False) -> "Batch":
        """Captures one launch sequence that builds every mesh into its tree (bvh_batch_create)."""
        k = len(dmeshes)
        if len(trees) != k:
            raise ValueError("one tree per mesh")
        hm = (C.c_void_p * k)(*[m.h.value if hasattr(m.h, "value") else m.h for m in dmeshes])
        ht = (C.c_void_p * k)(*[t.h.value if hasattr(t.h, "value") else t.h for t in trees])
        p = N.BuildParams(algo, radius, 0, radix_bits, N.BUILD_TUNE if tune else 0)
        h = C.c_void_p()
        self._check(self._lib.bvh_batch_create(self.h, hm, ht, k, C.byref(p), C.byref(h)), "batch create")
        return Batch(self, h, list(trees))

    def build_batched(self, dmeshes, trees, algo=ALGO_LBVH, radius=25, radix_bits=0) -> dict:
        """Eager segmented build of several meshes in one launch sequence (bvh_build_batched)."""
        k = len(dmeshes)
        hm = (C.c_void_p * k)(*[m.h.value if hasattr(m.h, "value") else m.h for m in dmeshes])
        ht = (C.c_void_p * k)(*[t.h.value if hasattr(t.h, "value") else t.h for t in trees])
        p = N.BuildParams(algo, radius, 0, radix_bits, 0)
        t = N.StageTimes()
        self._check(self._lib.bvh_build_batched(self.h, hm, ht, k, C.byref(p), C.byref(t)), "build_batched")
        return {"total_ms": float(t.total_ms), "stages": t.as_dict(), "ploc_iterations": int(t.ploc_iterations)}

    def build_async(self, dmesh: "DeviceMesh", tree: "Tree", algo=ALGO_LBVH, radius=25, rounds=1, radix_bits=0):
        """Enqueues a build of `tree` (bvh_build_async); the tree is valid after sync()."""
        p = N.BuildParams(algo, radius, rounds, radix_bits, 0)
        self._check(self._lib.bvh_build_async(self.h, dmesh.h, C.byref(p), tree.h), "build_async")

    def trace_async(self, tree: "Tree", dmesh: "DeviceMesh", rays: "DeviceBuffer", t: "DeviceBuffer",
                    prim: "DeviceBuffer", counters: "DeviceBuffer | None" = None):
        """Enqueues a closest-hit trace over device buffers (bvh_trace_async):
        rays 24 B, t 4 B, prim 4 B (and counters 12 B) per ray."""
        n = rays.nbytes // 24
        self._check(self._lib.bvh_trace_async(self.h, tree.h, dmesh.h, rays.ptr, n, t.ptr, prim.ptr,
                                              counters.ptr if counters is not None else None), "trace_async")

    def cache_info(self):
        """(bytes, blocks) of released device blocks cached for this context's device."""
        b, n = C.c_uint64(), C.c_uint32()
        self._check(self._lib.bvh_device_cache_info(self.device, C.byref(b), C.byref(n)), "cache info")
        return int(b.value), int(n.value)

    def cache_trim(self):
        """Frees the device's cached blocks (bvh_device_cache_trim)."""
        self._check(self._lib.bvh_device_cache_trim(self.device), "cache trim")

    def cache_set_limit(self, nbytes: int) -> int:
        """Most bytes the device's cache keeps (bvh_device_cache_set_limit); returns the previous limit."""
        prev = C.c_uint64()
        self._check(self._lib.bvh_device_cache_set_limit(self.device, int(nbytes), C.byref(prev)), "cache limit")
        return int(prev.value)

    def set_trace_order(self, order: int):
        """Visit order of this context's traces (bvh_ctx_set_trace_order):
        TRACE_REFERENCE (default: the reference renderer's order, ids, t and
        counters) or TRACE_NEAREST (nearer child first, least (t, index) hit)."""
        self._check(self._lib.bvh_ctx_set_trace_order(self.h, int(order)), "set_trace_order")

    def sync(self):
        """Waits for the context's queued work; raises the first failure of the
        asynchronous calls since the previous sync (bvh_sync)."""
        self._check(self._lib.bvh_sync(self.h), "sync")

    def buffer(self, nbytes: int) -> "DeviceBuffer":
        """Device memory of nbytes bytes (bvh_device_alloc)."""
        h = C.c_void_p()
        self._check(self._lib.bvh_device_alloc(self.h, max(16, nbytes), C.byref(h)), "device alloc")
        return DeviceBuffer(self, h.value, nbytes)

    def trace(self, tree: "Tree", dmesh: "DeviceMesh", rays: np.ndarray, counters=False):
        """Closest hit for host rays (n, 6) float32; returns (t, prim[, counters], ms)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 6)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        cnt = np.zeros((n, 3), np.int32) if counters else None
        ms = C.c_float(0)
        self._check(self._lib.bvh_trace_host(self.h, tree.h, dmesh.h, rays.ctypes.data, n, t.ctypes.data,
                                             prim.ctypes.data, cnt.ctypes.data if counters else None,
                                             C.byref(ms)), "trace")
        return (t, prim, cnt, float(ms.value)) if counters else (t, prim, float(ms.value))

    def debug_morton(self, dmesh: "DeviceMesh", algo=ALGO_LBVH) -> np.ndarray:
        out = np.zeros(dmesh.n, dtype=np.uint64)
        self._check(self._lib.bvh_debug_morton(self.h, dmesh.h, algo, out.ctypes.data), "debug_morton")
        return out

    def debug_sorted_keys(self, n: int):
        k = np.zeros(n, dtype=np.uint64)
        v = np.zeros(n, dtype=np.uint32)
        self._check(self._lib.bvh_debug_sorted_keys(self.h, k.ctypes.data, v.ctypes.data, n), "sorted keys")
        return k, v

    def sort_pairs(self, keys: np.ndarray, vals=None, begin_bit=0, end_bit=64, radix_bits=0):
        """Sorts host arrays on the device (stable LSD over [begin_bit, end_bit))."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        n = keys.size
        dk = C.c_void_p()
        self._check(self._lib.bvh_device_alloc(self.h, max(8, keys.nbytes), C.byref(dk)), "alloc")
        dv = C.c_void_p()
        try:
            self._check(self._lib.bvh_copy_h2d(self.h, dk, keys.ctypes.data, keys.nbytes), "h2d")
            if vals is not None:
                vals = np.ascontiguousarray(vals, dtype=np.uint32)
                self._check(self._lib.bvh_device_alloc(self.h, max(4, vals.nbytes), C.byref(dv)), "alloc")
                self._check(self._lib.bvh_copy_h2d(self.h, dv, vals.ctypes.data, vals.nbytes), "h2d")
            ms = C.c_float(0)
            self._check(self._lib.bvh_sort_pairs_u64(self.h, dk, dv if vals is not None else None, n, begin_bit,
                                                      end_bit, radix_bits, C.byref(ms)), "sort")
            ok = np.empty_like(keys)
            self._check(self._lib.bvh_copy_d2h(self.h, ok.ctypes.data, dk, keys.nbytes), "d2h")
            ov = None
            if vals is not None:
                ov = np.empty_like(vals)
                self._check(self._lib.bvh_copy_d2h(self.h, ov.ctypes.data, dv, vals.nbytes), "d2h")
            return ok, ov, float(ms.value)
        finally:
            self._lib.bvh_device_free(self.h, dk)
            if dv.value:
                self._lib.bvh_device_free(self.h, dv)


class Batch:
    """k meshes built into k trees by one captured hipGraph (bvh_batch_t)."""

    def __init__(self, ctx: Context, h, trees):
        self.ctx, self.h, self.trees = ctx, h, trees

    def run(self) -> float:
        ms = C.c_float(0)
        self.ctx._check(self.ctx._lib.bvh_batch_run(self.h, C.byref(ms)), "batch run")
        return float(ms.value)

    def info(self) -> dict:
        i = N.BatchInfo()
        self.ctx._check(self.ctx._lib.bvh_batch_get_info(self.h, C.byref(i)), "batch info")
        return i.as_dict()

    def free(self):
        if self.h:
            self.ctx._lib.bvh_batch_free(self.h)
            self.h = None


class DeviceBuffer:
    """A device allocation with synchronous host copies."""

    def __init__(self, ctx: Context, ptr: int, nbytes: int):
        self.ctx, self.ptr, self.nbytes = ctx, ptr, nbytes

    def write(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        self.ctx._check(self.ctx._lib.bvh_copy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "h2d")

    def read(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("read past the buffer")
        self.ctx._check(self.ctx._lib.bvh_copy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            self.ctx._lib.bvh_device_free(self.ctx.h, self.ptr)
            self.ptr = None


class DeviceMesh:
    def __init__(self, ctx: Context, h, n):
        self.ctx, self.h, self.n = ctx, h, n

    def free(self):
        if self.h:
            self.ctx._lib.bvh_mesh_free(self.ctx.h, self.h)
            self.h = None

    def release(self):
        """Returns the mesh's device blocks to the device cache (bvh_mesh_release):
        only when no queued work on any context uses the mesh."""
        if self.h:
            self.ctx._check(self.ctx._lib.bvh_mesh_release(self.ctx.h, self.h), "mesh release")
            self.h = None


class Tree:
    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def info(self):
        nn, npr = C.c_uint32(), C.c_uint32()
        self.ctx._lib.bvh_tree_info(self.h, C.byref(nn), C.byref(npr))
        return nn.value, npr.value

    def download(self):
        nn, npr = self.info()
        nodes = np.zeros(nn, dtype=NODE_DTYPE)
        idx = np.zeros(npr, dtype=np.uint32)
        self.ctx._check(self.ctx._lib.bvh_tree_download(self.ctx.h, self.h, nodes.ctypes.data, idx.ctypes.data),
                        "download")
        return nodes, idx

    def upload_nodes(self, nodes: np.ndarray):
        nodes = np.ascontiguousarray(nodes, dtype=NODE_DTYPE)
        self.ctx._check(self.ctx._lib.bvh_tree_from_nodes(self.ctx.h, nodes.ctypes.data, nodes.size, self.h),
                        "tree_from_nodes")

    def save(self, path: str):
        """Tree file (bvh_io.h): header, 48-byte BVHNode records, primitive indices."""
        self.ctx._check(N.host().bvh_tree_save(self.ctx.h, self.h, path.encode()), "tree save")

    def load(self, path: str):
        self.ctx._check(N.host().bvh_tree_load(self.ctx.h, path.encode(), self.h), "tree load")

    def export_viz(self, path: str):
        """The reference's 32-byte VizNode export (exportBVHToBinary)."""
        self.ctx._check(N.host().bvh_tree_export_viz(self.ctx.h, self.h, path.encode()), "tree export")

    def free(self):
        if self.h:
            self.ctx._lib.bvh_tree_free(self.ctx.h, self.h)
            self.h = None

    def release(self):
        """bvh_tree_release: like release() of a mesh."""
        if self.h:
            self.ctx._check(self.ctx._lib.bvh_tree_release(self.ctx.h, self.h), "tree release")
            self.h = None


class _Builder:
    """Mirror of the reference BVHBuilder (include/bvh_builder.h:10-37)."""
    algo = ALGO_LBVH

    def __init__(self, ctx: Context | None = None):
        self._ctx = ctx
        self._nodes = np.zeros(0, dtype=NODE_DTYPE)
        self._indices = np.zeros(0, dtype=np.uint32)
        self._times = {"total_ms": 0.0, "stages": {}}
        self._params = {}

    def _context(self):
        if self._ctx is None:
            self._ctx = Context(0)
        return self._ctx

    def build(self, mesh: Mesh) -> None:
        ctx = self._context()
        dm = ctx.upload(mesh)
        tree = ctx.tree()
        try:
            self._times = ctx.build(dm, tree, algo=self.algo, **self._params)
            self._nodes, self._indices = tree.download()
        finally:
            tree.free()
            dm.free()

    def get_nodes(self):
        return self._nodes

    def get_indices(self):
        return self._indices

    def get_last_build_time_ms(self) -> float:
        return self._times["total_ms"]

    def _stage_sum(self, prefix):
        return sum(v for k, v in self._times["stages"].items() if k.startswith(prefix))

    def get_timing_breakdown(self) -> str:
        """The reference's stage lines (lbvh_builder.cu:490-498 and the
        LBVH+/PLOC counterparts); fused stages read 0 (the C++ classes give
        the same text).  The finer per-kernel stages are in `stages()`."""
        s = self._stage_sum
        if self.algo == ALGO_LBVH:
            rows = [("Bounds/Centroids", s("Bounds/Centroids")), ("Morton Codes", s("Morton Codes")),
                    ("CUB Radix Sort", s("Sort Pass")), ("Topology Build", s("Topology+Refit")),
                    ("Bottom-Up Refit", 0.0)]
        elif self.algo == ALGO_LBVH_PLUS:
            opt = s("Treelet Round")
            rows = [("Bounds/Centroids", s("Bounds/Centroids")), ("Morton Codes", s("Morton Codes")),
                    ("Radix Sort", s("Sort Pass")), ("Topology Build", s("Topology+Refit")),
                    ("Refit + Optimize", opt)]
        else:
            rows = [("Initialization", s("Bounds/Centroids") + s("Morton Codes") + s("Sort Pass") + s("PLOC Init")),
                    ("NN Search (Red)", s("PLOC Global") + s("PLOC LDS Tail")), ("Merging (Green)", 0.0),
                    ("Compaction (Mag)", 0.0)]
        lines = [f"  {name + ':':<18}{v:.3f} ms" for name, v in rows]
        if self.algo == ALGO_LBVH_PLUS:
            lines[-1] += f" (opt: {rows[-1][1]:.3f} ms)"
        return "\n".join(lines)

    def stages(self) -> dict:
        return dict(self._times["stages"])

    def cleanup(self):
        self._nodes = np.zeros(0, dtype=NODE_DTYPE)
        self._indices = np.zeros(0, dtype=np.uint32)


class LBVHBuilder(_Builder):
    algo = ALGO_LBVH

    def get_name(self):
        return "LBVH"


class LBVHPlusBuilder(_Builder):
    algo = ALGO_LBVH_PLUS

    def __init__(self, ctx=None, rounds: int = 1):
        super().__init__(ctx)
        self._params = {"rounds": rounds}

    def get_name(self):
        return "LBVH+"


class PLOCBuilder(_Builder):
    algo = ALGO_PLOC

    def __init__(self, radius: int = 25, ctx=None):
        super().__init__(ctx)
        self.radius = radius
        self._params = {"radius": radius}

    def get_name(self):
        return f"PLOC CUDA (r={self.radius})"
