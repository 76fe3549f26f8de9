"""Serving loop: Python side of include/bvh_server.h.

`Server` wraps the native continuous-batching server (libbvhhost.so): slots
with their own HIP streams take FIFO groups of build+trace requests that fit
the device-memory budget, build every new mesh of a group in one segmented
launch sequence, and trace each request's rays against its own tree.

The open-loop driver (`run_open_loop`) submits requests at Poisson arrival
times and reads the server's own timestamps, so latency = done - arrival is
measured on the server clock.  `run_open_loop` and `latency_stats` take any
object with the Server methods, which is how the CPU tests drive them with a
synthetic clock.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _native as N

ALGO_LBVH, ALGO_PLOC = N.ALGO_LBVH, N.ALGO_PLOC


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("max_group", C.c_int32),
                ("max_group_tris", C.c_uint32), ("memory_budget", C.c_uint64), ("algo", C.c_int32),
                ("ploc_radius", C.c_int32), ("trace_order", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("mesh_key", C.c_uint64), ("soa", C.c_void_p * 9), ("ntri", C.c_uint32),
                ("nrays", C.c_uint32), ("rays", C.c_void_p), ("t_out", C.c_void_p), ("prim_out", C.c_void_p),
                ("arrival", C.c_double)]


class Completion(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int32), ("slot", C.c_int32), ("group_size", C.c_uint32),
                ("built", C.c_uint32), ("arrival", C.c_double), ("admitted", C.c_double), ("done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("budget", C.c_uint64), ("in_use", C.c_uint64), ("peak", C.c_uint64), ("resident_bytes", C.c_uint64),
                ("resident_meshes", C.c_uint32), ("queued", C.c_uint32), ("in_flight", C.c_uint32),
                ("groups", C.c_uint32), ("completed", C.c_uint32), ("evictions", C.c_uint32),
                ("rejected", C.c_uint32), ("max_group_seen", C.c_uint32)]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_SIGS = {
    "bvh_server_create": (C.c_int32, [C.POINTER(ServerConfig), C.POINTER(C.c_void_p)]),
    "bvh_server_submit": (C.c_int32, [C.c_void_p, C.POINTER(Request)]),
    "bvh_server_poll": (C.c_int32, [C.c_void_p, C.POINTER(Completion), C.c_uint32, C.POINTER(C.c_uint32),
                                    C.c_int32]),
    "bvh_server_now": (C.c_double, [C.c_void_p]),
    "bvh_server_get_stats": (C.c_int32, [C.c_void_p, C.POINTER(ServerStats)]),
    "bvh_server_drain": (C.c_int32, [C.c_void_p, C.c_int32]),
    "bvh_server_last_error": (C.c_char_p, [C.c_void_p]),
    "bvh_server_destroy": (C.c_int32, [C.c_void_p]),
    "bvh_serve_footprint": (None, [C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_uint64)]),
    "bvh_serve_form_group": (C.c_uint32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64, C.c_uint32,
                                          C.c_uint32]),
}
_bound = False


def lib():
    global _bound
    L = N.host()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def footprint(ntri: int, nrays: int, algo: int = ALGO_LBVH):
    tr, rs = C.c_uint64(), C.c_uint64()
    lib().bvh_serve_footprint(ntri, nrays, algo, C.byref(tr), C.byref(rs))
    return int(tr.value), int(rs.value)


def form_group(bytes_, tris, runnable, free_bytes, max_group, max_group_tris) -> int:
    b = np.ascontiguousarray(bytes_, np.uint64)
    t = np.ascontiguousarray(tris, np.uint32)
    r = np.ascontiguousarray(runnable, np.uint8)
    return int(lib().bvh_serve_form_group(b.ctypes.data, t.ctypes.data, r.ctypes.data, b.size, free_bytes,
                                          max_group, max_group_tris))


class Server:
    """Native server: `slots` streams, groups of up to `max_group` requests."""

    def __init__(self, algo=ALGO_LBVH, radius=16, slots=4, max_group=32, max_group_tris=0, budget=0, device=0,
                 trace_order=0):
        L = lib()
        cfg = ServerConfig(device, slots, max_group, max_group_tris, budget, algo, radius, trace_order)
        h = C.c_void_p()
        st = L.bvh_server_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise N.NativeError(st, "bvh_server_create")
        self.h = h
        self._keep = {}  # request id -> host buffers that must outlive the request

    def now(self) -> float:
        return float(lib().bvh_server_now(self.h))

    def wait_until(self, t: float):
        while True:
            d = t - self.now()
            if d <= 0:
                return
            time.sleep(min(d, 0.002) if d > 0.0005 else 0)

    def submit(self, rid, mesh=None, rays=None, key=0, arrival=None):
        """mesh: bvh_amd.Mesh or None (trace against resident `key`); rays: (n, 6) float32.
        Returns (t, prim) arrays that are filled once the request completes."""
        r = Request()
        r.id = rid
        r.mesh_key = key
        keep = []
        if mesh is not None:
            soa = np.ascontiguousarray(mesh.soa, np.float32)
            keep.append(soa)
            for k in range(9):
                r.soa[k] = soa[k].ctypes.data
            r.ntri = soa.shape[1]
        rays = np.ascontiguousarray(rays if rays is not None else np.zeros((0, 6)), np.float32).reshape(-1, 6)
        n = rays.shape[0]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        keep += [rays, t, prim]
        r.nrays = n
        r.rays = rays.ctypes.data if n else None
        r.t_out = t.ctypes.data if n else None
        r.prim_out = prim.ctypes.data if n else None
        r.arrival = self.now() if arrival is None else arrival
        self._keep[rid] = keep
        st = lib().bvh_server_submit(self.h, C.byref(r))
        if st != 0:
            del self._keep[rid]
            raise N.NativeError(st, "bvh_server_submit")
        return t, prim

    def poll(self, wait_ms=0, max_n=256):
        out = (Completion * max_n)()
        cnt = C.c_uint32()
        lib().bvh_server_poll(self.h, out, max_n, C.byref(cnt), wait_ms)
        res = []
        for i in range(cnt.value):
            c = out[i]
            self._keep.pop(c.id, None)
            res.append({"id": int(c.id), "status": int(c.status), "slot": int(c.slot),
                        "group_size": int(c.group_size), "built": int(c.built), "arrival": float(c.arrival),
                        "admitted": float(c.admitted), "done": float(c.done)})
        return res

    def drain(self, timeout_ms=600_000):
        st = lib().bvh_server_drain(self.h, timeout_ms)
        if st != 0:
            raise N.NativeError(st, "bvh_server_drain timed out")

    def stats(self) -> dict:
        s = ServerStats()
        lib().bvh_server_get_stats(self.h, C.byref(s))
        return s.as_dict()

    def last_error(self) -> str:
        return lib().bvh_server_last_error(self.h).decode()

    def close(self):
        if getattr(self, "h", None):
            lib().bvh_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---- workload ------------------------------------------------------------------

def log_uniform_ints(rng, n, lo, hi):
    """floor(exp(ln lo + u ln(hi/lo))), u ~ U[0,1)."""
    u = rng.random(n)
    return np.floor(np.exp(math.log(lo) + u * math.log(hi / lo))).astype(np.int64)


def poisson_arrivals(rate_per_s: float, count: int, seed: int) -> np.ndarray:
    """Arrival offsets (s) of a Poisson process: cumulative Exp(1/rate) gaps, first at 0."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, count)
    gaps[0] = 0.0
    return np.cumsum(gaps)


def make_requests(seed, count, tri_lo=10_000, tri_hi=2_000_000, ray_lo=65_536, ray_hi=1_048_576,
                  follow_up=0.2, pool=None):
    """Build+trace requests (blob meshes, rays from the bounding sphere toward
    the box) and a fraction of follow-up trace requests on an earlier mesh.

    pool: if set, only `pool` distinct (mesh, rays) pairs are generated and
    each request takes one of them at random; every build request still
    builds its own copy under its own key (host memory and generation time
    stay bounded for thousands of requests)."""
    import bvh_amd as B
    rng = np.random.default_rng(seed)
    distinct = count if pool is None else min(pool, count)
    tris = log_uniform_ints(rng, distinct, tri_lo, tri_hi)
    nrays = log_uniform_ints(rng, distinct, ray_lo, ray_hi)
    follow = rng.random(count) < follow_up
    pick = np.arange(count) if pool is None else rng.integers(0, distinct, count)
    meshes, boxes_p, rays_p = {}, {}, {}

    def entry(p):
        if p not in meshes:
            m = B.Mesh.blobs(seed * 7919 + int(p), int(tris[p]))
            meshes[p] = m
            boxes_p[p] = m.bounds()
            rays_p[p] = B.rays_sphere_to_box(seed * 1000003 + int(p), boxes_p[p], int(nrays[p]))
        return meshes[p], boxes_p[p], rays_p[p]

    reqs = []
    built = {}  # key -> pool entry of its mesh
    for i in range(count):
        p = int(pick[i])
        if follow[i] and built:
            keys = sorted(built)
            key = keys[int(rng.integers(len(keys)))]
            if pool is None:
                rays = B.rays_sphere_to_box(seed * 1000003 + i, boxes_p[built[key]], int(nrays[p]))
            else:
                rays = rays_p[built[key]]
            reqs.append({"id": i, "key": key, "mesh": None, "ntri": 0, "rays": rays})
            continue
        m, _, rays = entry(p)
        key = i + 1
        built[key] = p
        reqs.append({"id": i, "key": key, "mesh": m, "ntri": m.n, "rays": rays})
    return reqs


# ---- driver and statistics ----------------------------------------------------------

def run_open_loop(server, reqs, offsets, lead=0.02):
    """Submits reqs[i] at server time t0 + offsets[i] and waits for all of
    them; returns (completions in id order, outputs by id, t0)."""
    t0 = server.now() + lead
    outs = {}
    done = []
    for q, off in zip(reqs, offsets):
        server.wait_until(t0 + float(off))
        outs[q["id"]] = server.submit(q["id"], q["mesh"], q["rays"], key=q["key"], arrival=t0 + float(off))
        done += server.poll(0)
    while len(done) < len(reqs):
        got = server.poll(200)
        done += got
    done.sort(key=lambda c: c["id"])
    return done, outs, t0


def percentile_nearest_rank(values, q):
    v = np.sort(np.asarray(values, np.float64))
    if v.size == 0:
        return float("nan")
    k = max(1, int(math.ceil(q * v.size)))
    return float(v[k - 1])


def latency_stats(done, t0=None):
    lat = np.array([c["done"] - c["arrival"] for c in done], np.float64)
    start = min(c["arrival"] for c in done) if t0 is None else t0
    span = max(c["done"] for c in done) - start
    return {"requests": len(done), "p50_ms": round(percentile_nearest_rank(lat, 0.50) * 1e3, 3),
            "p99_ms": round(percentile_nearest_rank(lat, 0.99) * 1e3, 3),
            "mean_ms": round(float(lat.mean()) * 1e3, 3),
            "throughput_rps": round(len(done) / span, 2) if span > 0 else None,
            "failed": int(sum(1 for c in done if c["status"] != 0))}
