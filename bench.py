#!/usr/bin/env python3
"""Benchmark of the MI355X BVH engine.

Default workload: LBVH build of 4,000,000 fp32 blob triangles (seed 7; 64
Gaussian blobs + 5% large triangles, shuffled), mesh resident in HBM.  A step
is one complete build (bounds, Morton, radix sort, topology + refit) of that
mesh.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The CPU baseline's OpenMP threads stay on neighbouring cores (read when the
# oracle's OpenMP runtime starts).
os.environ.setdefault("OMP_PROC_BIND", "close")

METRIC = "build Mtri/s @4M tris; batched vs plain meshes/s; p50/p99 build+trace latency"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of each LBVH stage (per triangle, or per key for a
# sort pass).  Bounds: 9 fp32 vertex words in + one 32-B box out.  Morton: box
# in + 8-B key out.  Sort pass: 8-B key in + out.  Topology+refit: the
# reference's hierarchy priced as the survey of it does (Karras: keys 8 +
# children 8 + parents 8; leaves: key 8 + vertex gather 36 + node 48; refit:
# parent and counter traffic, children, sibling box and node write 84), 200
# B/tri.  This engine's level hierarchy needs less (key 8 + leaf box gather 32
# + leaf node 48 + internal node 48 + index 4 = 140 B/tri); the roofline
# reports the launch at that figure too.
STAGE_BYTES = {"Bounds/Centroids": 68, "Morton Codes": 40, "Sort Pass": 16, "Topology+Refit": 200}
DESIGN_HIER_BYTES = 140


# HBM bytes per dispatch of the dominant kernels, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) of this workload, scaled by the same counters on a
# 256 MiB copy kernel; the summary and the commands that produced it are in
# profiles/.
TRAFFIC_FILE_REL = "profiles/traffic_lbvh_4m_blobs.json"
TRAFFIC_FILE = os.path.join(ROOT, TRAFFIC_FILE_REL)


# Algorithmic bytes of the companion workloads: a treelet round moves about
# (13 node reads + 6 node writes) x 48 B per internal node; a ray 24 (ray) + 8
# (hit) + 48 per node visited + 36 per triangle tested, counted by an
# instrumented run; PLOC 124 B per cluster per iteration plus 48 B per new
# node, with the clusters per iteration from profiles/ploc_iterations_*.json.
TREELET_BYTES_PER_NODE = 912


def ploc_bytes(name, n, radius):
    try:
        with open(os.path.join(ROOT, "profiles", name)) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("triangles") != n or d.get("radius") != radius:
        return None
    return int(d["algorithmic_bytes"])


def roof(bytes_, ms):
    if not bytes_ or not ms:
        return None
    gbs = bytes_ / (ms * 1e-3) / 1e9
    return {"bytes": int(bytes_), "ms": round(ms, 4), "GBps": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}


def pmc_traffic(kernel, n):
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("triangles") != n:
        return None
    k = d.get("kernels", {}).get(kernel)
    return int(k["hbm_bytes_per_dispatch"]) if k else None


def stage_unit_bytes(name):
    for k, v in STAGE_BYTES.items():
        if name.startswith(k):
            return v
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_builds(ctx, dm, tree, steps, warmup, **kw):
    """Back-to-back synchronous builds of a resident mesh.  The wall-clock loop
    records no per-stage events; a second loop of the same length records them
    for the mean device time and per-stage times."""
    for _ in range(warmup):
        ctx.build(dm, tree, **kw)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.build(dm, tree, timing=False, **kw)
    wall = time.perf_counter() - t0
    stage_sum = {}
    dev_total = 0.0
    iters = 0
    for _ in range(steps):
        t = ctx.build(dm, tree, **kw)
        dev_total += t["total_ms"]
        iters = t["ploc_iterations"]
        for k, v in t["stages"].items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    return wall, dev_total / steps, {k: v / steps for k, v in stage_sum.items()}, iters


def extras(ctx, dm, tree, mesh, steps):
    """Config-1 companions of the headline: LBVH+ (3 rounds), PLOC (r=25), and
    1024x768 camera rays traced through the LBVH and PLOC trees."""
    import bvh_amd as B
    out = {}
    for name, kw in (("lbvh_plus_3_rounds", {"algo": B.ALGO_LBVH_PLUS, "rounds": 3}),
                     ("ploc_r25", {"algo": B.ALGO_PLOC, "radius": 25})):
        log(f"[bench] {name}")
        wall, dev, st, iters = timed_builds(ctx, dm, tree, max(3, steps // 4), 1, **kw)
        n_steps = max(3, steps // 4)
        out[name] = {"ms_per_build": round(wall / n_steps * 1e3, 3), "device_ms": round(dev, 3),
                     "Mtri_s": round(mesh.n / (wall / n_steps) / 1e6, 1),
                     "stages_ms": {k: round(v, 4) for k, v in st.items()}}
        if iters:
            out[name]["iterations"] = iters
        if name.startswith("lbvh_plus"):
            tl = sum(v for k, v in st.items() if k.startswith("Treelet Round"))
            out[name]["treelet_roofline"] = roof(TREELET_BYTES_PER_NODE * (mesh.n - 1) * 3, tl)
        else:
            loop = st.get("PLOC Global Iterations", 0.0) + st.get("PLOC LDS Tail", 0.0)
            out[name]["iterations_roofline"] = roof(ploc_bytes("ploc_iterations_4m_blobs_r25.json", mesh.n, 25), loop)
    rays = B.rays_sphere_camera(mesh.bounds(), 1024, 768)
    for name, kw in (("trace_lbvh", {"algo": B.ALGO_LBVH}), ("trace_ploc_r25", {"algo": B.ALGO_PLOC, "radius": 25})):
        log(f"[bench] {name}")
        ctx.build(dm, tree, **kw)
        out[name], ref = trace_stats(ctx, tree, dm, rays, 5)
        log(f"[bench] {name} (nearest-first order)")
        ctx.set_trace_order(B.TRACE_NEAREST)
        try:
            out[name + "_nearest"], near = trace_stats(ctx, tree, dm, rays, 5)
        finally:
            ctx.set_trace_order(B.TRACE_REFERENCE)
        out[name + "_nearest"].update(same_hits(ref, near))
    return out


def trace_stats(ctx, tree, dm, rays, reps):
    """One counted trace, then the median device time of `reps` traces; the
    algorithmic bytes per ray are 24 (ray) + 8 (hit) + 48 per node visited +
    36 per triangle tested."""
    t, prim, cnt, _ = ctx.trace(tree, dm, rays, counters=True)
    ms = float(np.median([ctx.trace(tree, dm, rays)[2] for _ in range(reps)]))
    per_ray = 24 + 8 + 48 * float(cnt[:, 0].mean()) + 36 * float(cnt[:, 2].mean())
    return ({"rays": int(rays.shape[0]), "device_ms": round(ms, 4),
             "Mrays_s": round(rays.shape[0] / ms / 1e3, 1), "hit_fraction": round(float((prim >= 0).mean()), 4),
             "avg_nodes_visited": round(float(cnt[:, 0].mean()), 2),
             "avg_tri_tests": round(float(cnt[:, 2].mean()), 2),
             "roofline": roof(per_ray * rays.shape[0], ms)}, (t, prim))


def same_hits(ref, near):
    """How the nearest-first hits compare with the reference order's."""
    (rt, rp), (nt, np_) = ref, near
    hit = rp >= 0
    rel = np.abs(nt[hit].astype(np.float64) - rt[hit]) / np.maximum(np.abs(rt[hit].astype(np.float64)), 1e-30)
    return {"rays_with_other_hit_than_reference_order": int((rp != np_).sum()),
            "max_rel_t_difference": float(rel.max()) if rel.size else 0.0}


def batched_vs_plain(ctx, steps, k=512, lo=1000, hi=200_000, seed=11):
    """512 uneven meshes (log-uniform 1,000-200,000 triangles, seed 11, each a
    uniform-cube mesh with its own seed, scale and offset): one captured
    launch sequence per algorithm against 512 plain per-mesh builds, meshes/s
    of each, and per-mesh bit-identity of the trees."""
    import bvh_amd as B
    counts = B.batch_counts(seed, k, lo, hi)
    log(f"[bench] batched: generating {k} meshes, {int(counts.sum()):,} triangles")
    meshes = [B.Mesh.batch_mesh(seed, i, int(c)) for i, c in enumerate(counts)]
    dms = [ctx.upload(m) for m in meshes]
    del meshes
    out = {"meshes": k, "triangles": int(counts.sum()), "seed": seed}
    for name, algo, radius in (("lbvh", B.ALGO_LBVH, 25), ("ploc_r16", B.ALGO_PLOC, 16)):
        log(f"[bench] batched {name}: plain per-mesh builds")
        plain_trees = [ctx.tree() for _ in dms]
        for dm, t in zip(dms, plain_trees):  # warm (allocations)
            ctx.build(dm, t, algo=algo, radius=radius)
        reps = max(1, min(3, steps // 5))
        t0 = time.perf_counter()
        for _ in range(reps):
            for dm, t in zip(dms, plain_trees):
                ctx.build(dm, t, algo=algo, radius=radius)
        plain_s = (time.perf_counter() - t0) / reps
        log(f"[bench] batched {name}: capture + tune")
        trees = [ctx.tree() for _ in dms]
        batch = ctx.batch(dms, trees, algo=algo, radius=radius, tune=True)
        batch.run()
        dev = []
        t0 = time.perf_counter()
        for _ in range(steps):
            dev.append(batch.run())
        batch_s = (time.perf_counter() - t0) / steps
        info = batch.info()
        log(f"[bench] batched {name}: bit-identity check")
        identical = True
        for a, b in zip(trees, plain_trees):
            na, ia = a.download()
            nb, ib = b.download()
            if na.tobytes() != nb.tobytes() or not np.array_equal(ia, ib):
                identical = False
                break
        batch.free()
        for t in trees + plain_trees:
            t.free()
        out[name] = {
            "plain_ms": round(plain_s * 1e3, 2), "plain_meshes_s": round(k / plain_s, 1),
            "batched_ms": round(batch_s * 1e3, 3), "batched_device_ms": round(float(np.median(dev)), 3),
            "batched_meshes_s": round(k / batch_s, 1), "speedup": round(plain_s / batch_s, 2),
            "batched_Mtri_s": round(int(counts.sum()) / batch_s / 1e6, 1),
            "bit_identical_all_meshes": identical,
            "graph_nodes": info["graph_nodes"], "radix_bits": info["radix_bits"], "sort_passes": info["sort_passes"],
            "key_bits": info["key_bits"], "tune": info["tune"],
        }
        if algo == B.ALGO_PLOC:
            out[name]["global_iterations"] = info["ploc_iterations"]
            out[name]["nn_grid"] = info["ploc_nn_grid"]
    for d in dms:
        d.free()
    return out


def batched_eager(ctx, steps, k=512, lo=1000, hi=200_000, seed=11):
    """The same 512 meshes built by the eager segmented path
    (bvh_build_batched: the batch's kernels launched directly, no captured
    graph), for kernel traces and counters of the segmented kernels."""
    import bvh_amd as B
    counts = B.batch_counts(seed, k, lo, hi)
    log(f"[bench] batched (eager): generating {k} meshes, {int(counts.sum()):,} triangles")
    meshes = [B.Mesh.batch_mesh(seed, i, int(c)) for i, c in enumerate(counts)]
    dms = [ctx.upload(m) for m in meshes]
    del meshes
    out = {"meshes": k, "triangles": int(counts.sum()), "seed": seed}
    for name, algo, radius in (("lbvh", B.ALGO_LBVH, 25), ("ploc_r16", B.ALGO_PLOC, 16)):
        log(f"[bench] batched (eager) {name}")
        trees = [ctx.tree() for _ in dms]
        ctx.build_batched(dms, trees, algo=algo, radius=radius)
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.build_batched(dms, trees, algo=algo, radius=radius)
        ms = (time.perf_counter() - t0) / steps * 1e3
        for t in trees:
            t.free()
        out[name] = {"ms": round(ms, 3), "meshes_s": round(k / ms * 1e3, 1),
                     "Mtri_s": round(int(counts.sum()) / ms / 1e3, 1)}
    for d in dms:
        d.free()
    return out


def large_strands(ctx, steps=2):
    """32,000,000 triangles: 200,000 random-walk strands x 160 segments (step
    0.01, seed 99); PLOC r=100 and LBVH+ (3 rounds) builds, then 4,194,304
    incoherent rays (uniform origins in the cube, uniform directions) traced
    through each tree."""
    import bvh_amd as B
    log("[bench] large: generating 200,000 strands x 160 segments (seed 99)")
    mesh = B.Mesh.strands(99, 200_000, 160)
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    rays = B.rays_uniform_cube(99, 4_194_304)
    out = {"triangles": mesh.n, "rays": int(rays.shape[0])}
    for name, kw in (("ploc_r100", {"algo": B.ALGO_PLOC, "radius": 100}),
                     ("lbvh_plus_3_rounds", {"algo": B.ALGO_LBVH_PLUS, "rounds": 3})):
        log(f"[bench] large: {name}")
        wall, dev, st, iters = timed_builds(ctx, dm, tree, steps, 1, **kw)
        r = {"ms_per_build": round(wall / steps * 1e3, 2), "device_ms": round(dev, 2),
             "Mtri_s": round(mesh.n / (wall / steps) / 1e6, 1), "stages_ms": {k: round(v, 3) for k, v in st.items()}}
        if iters:
            r["iterations"] = iters
        if name.startswith("lbvh_plus"):
            tl = sum(v for k, v in st.items() if k.startswith("Treelet Round"))
            r["treelet_roofline"] = roof(TREELET_BYTES_PER_NODE * (mesh.n - 1) * 3, tl)
        else:
            loop = st.get("PLOC Global Iterations", 0.0) + st.get("PLOC LDS Tail", 0.0)
            r["iterations_roofline"] = roof(ploc_bytes("ploc_iterations_32m_strands_r100.json", mesh.n, 100), loop)
        log(f"[bench] large: trace through {name}")
        st_ref, ref = trace_stats(ctx, tree, dm, rays, 1)
        r["trace_device_ms"] = round(st_ref["device_ms"], 2)
        r["trace_Mrays_s"] = st_ref["Mrays_s"]
        r["hit_fraction"] = st_ref["hit_fraction"]
        r["trace_roofline"] = st_ref["roofline"]
        log(f"[bench] large: trace through {name} (nearest-first order)")
        ctx.set_trace_order(B.TRACE_NEAREST)
        try:
            st_near, near = trace_stats(ctx, tree, dm, rays, 1)
        finally:
            ctx.set_trace_order(B.TRACE_REFERENCE)
        st_near.update(same_hits(ref, near))
        r["trace_nearest"] = st_near
        out[name] = r
    tree.free()
    dm.free()
    return out


SERVERS = {"coalesced": {"slots": 4, "max_group": 32}, "plain": {"slots": 1, "max_group": 1},
           "coalesced_nearest": {"slots": 4, "max_group": 32, "trace_order": 1}}


def serve_child(spec):
    """One server, one measurement (run in a process of its own): the request
    stream, closed-loop (rate None) or Poisson at `rate` requests/s; prints
    one JSON line."""
    from bvh_amd import serving as S
    count, seed, pool = spec["count"], spec["seed"], spec["pool"]
    reqs = S.make_requests(seed, count, pool=pool)
    warm = S.make_requests(seed + 1, 4, tri_lo=10_000, tri_hi=50_000, ray_lo=65_536, ray_hi=100_000, follow_up=0.0)
    for q in warm:
        q["key"] = 0
    rate = spec["rate"]
    offsets = np.zeros(count) if rate is None else S.poisson_arrivals(rate, count, seed)
    with S.Server(algo=S.ALGO_LBVH, **SERVERS[spec["server"]]) as srv:
        S.run_open_loop(srv, warm, np.zeros(len(warm)))
        done, _, t0 = S.run_open_loop(srv, reqs, offsets)
        st = S.latency_stats(done, t0)
        ss = srv.stats()
    st["groups"] = ss["groups"]
    st["max_group"] = ss["max_group_seen"]
    st["evictions"] = ss["evictions"]
    st["peak_device_GB"] = round(ss["peak"] / 1e9, 2)
    if rate is not None:
        st["offered_rps"] = round(rate, 2)
    st["triangles_total"] = int(sum(q["ntri"] for q in reqs))
    st["rays_total"] = int(sum(q["rays"].shape[0] for q in reqs))
    print(json.dumps(st), flush=True)


def serving_latency(count=1000, seed=23, pool=96):
    """Build+trace requests (blob meshes log-uniform 10K-2M triangles, rays
    log-uniform 65,536-1,048,576 from the bounding sphere toward the box, 20%
    follow-up traces on resident meshes), LBVH builds; `pool` distinct
    (mesh, rays) pairs drawn at random by the `count` requests, each build
    request under its own key.  Capacity = closed-loop saturation (all
    requests at t=0) of each server; then Poisson arrivals at 50% and 90% of
    each server's own capacity, and the first two at 90% of the plain path's
    capacity (the same offered load).  Servers: coalescing (4 slots, groups
    of up to 32), plain (1 slot, one mesh per launch) and coalescing with the
    nearest-first trace order; p50/p99 of done - arrival on the server clock.
    Each measurement runs in a fresh process, as a deployed server does (a
    server started in a process where another one has already served
    measures slower)."""
    import subprocess

    def one(name, rate):
        spec = {"server": name, "rate": rate, "count": count, "seed": seed, "pool": pool}
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--serve-child", json.dumps(spec)],
                           stdout=subprocess.PIPE, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"serving measurement {spec} failed with status {p.returncode}")
        return json.loads(p.stdout.strip().splitlines()[-1])

    out = {"requests": count, "distinct_meshes": pool, "seed": seed, "algo": "LBVH",
           "process_per_measurement": True}
    for name in SERVERS:
        log(f"[bench] serving: capacity ({name})")
        r = one(name, None)
        out["triangles_total"] = r.pop("triangles_total")
        out["rays_total"] = r.pop("rays_total")
        out[f"capacity_{name}"] = r
    for name in SERVERS:
        cap = out[f"capacity_{name}"]["throughput_rps"]
        for load in (0.5, 0.9):
            log(f"[bench] serving: {int(load * 100)}% of its capacity ({name})")
            r = one(name, load * cap)
            r.pop("triangles_total")
            r.pop("rays_total")
            out[f"load{int(load * 100)}_{name}"] = r
    same = 0.9 * out["capacity_plain"]["throughput_rps"]
    log("[bench] serving: coalescing at the plain path's 90% load")
    r = one("coalesced", same)
    r.pop("triangles_total")
    r.pop("rays_total")
    out["plain90_load_coalesced"] = r
    return out


def cpu_model():
    """The host CPU model and logical CPU count (from /proc/cpuinfo)."""
    model, count = None, 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    count += 1
                    model = model or line.split(":", 1)[1].strip()
    except OSError:
        pass
    return {"model": model, "logical_cpus": count or os.cpu_count()}


def _median_time(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    times, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), out


def cpu_baseline(mesh, gpu_nodes, gpu_idx, sample=1_000_000, serial_sample=250_000):
    """The oracle on this host's cores (OMP_PROC_BIND=close): the headline
    OpenMP LBVH build of the full mesh, median of 5 after 1 warmup; OpenMP
    LBVH+ (3 rounds) and PLOC (r=25) on the first `sample` triangles of the
    same (shuffled) mesh and OpenMP traversal of every 12th of the 1024x768
    camera rays through the full-mesh LBVH, median of 5; serial figures on
    smaller samples (every 96th ray), median of 3."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    import bvh_amd as B
    threads = O.lib().orc_omp_threads()
    med, (nodes, idx, _) = _median_time(lambda: O.build_lbvh_omp(mesh.soa), 5)
    parity = (O.nodes_equal(gpu_nodes, nodes) is None) and bool(np.array_equal(gpu_idx, idx))
    sub = np.ascontiguousarray(mesh.soa[:, :sample])
    small = np.ascontiguousarray(mesh.soa[:, :serial_sample])
    extra = {}
    m, _ = _median_time(lambda: O.build_lbvh_plus_omp(sub, 3), 5)
    extra["lbvh_plus_3_rounds_omp"] = {"triangles": sample, "ms": round(m * 1e3, 1), "Mtri_s": round(sample / m / 1e6, 3)}
    m, r = _median_time(lambda: O.build_ploc_omp(sub, 25), 5)
    extra["ploc_r25_omp"] = {"triangles": sample, "ms": round(m * 1e3, 1), "Mtri_s": round(sample / m / 1e6, 3),
                             "iterations": int(len(r[3]))}
    rays = B.rays_sphere_camera(mesh.bounds(), 1024, 768)
    ro = np.ascontiguousarray(rays[::12])  # 65,536 rays spread over the image
    m, _ = _median_time(lambda: O.trace(nodes, mesh.soa, ro, omp=True), 5)
    extra["trace_lbvh_omp"] = {"rays": int(ro.shape[0]), "ms": round(m * 1e3, 1),
                               "Mrays_s": round(ro.shape[0] / m / 1e6, 4)}
    serial = {}
    m, _ = _median_time(lambda: O.build_lbvh(sub), 3)
    serial["lbvh"] = {"triangles": sample, "ms": round(m * 1e3, 1), "Mtri_s": round(sample / m / 1e6, 3)}
    m, _ = _median_time(lambda: O.build_lbvh_plus(small, 3), 3)
    serial["lbvh_plus_3_rounds"] = {"triangles": serial_sample, "ms": round(m * 1e3, 1),
                                    "Mtri_s": round(serial_sample / m / 1e6, 3)}
    m, _ = _median_time(lambda: O.build_ploc(small, 25), 3)
    serial["ploc_r25"] = {"triangles": serial_sample, "ms": round(m * 1e3, 1),
                          "Mtri_s": round(serial_sample / m / 1e6, 3)}
    rs = np.ascontiguousarray(rays[::96])  # 8,192 rays spread over the image
    m, _ = _median_time(lambda: O.trace(nodes, mesh.soa, rs), 3)
    serial["trace_lbvh"] = {"rays": int(rs.shape[0]), "ms": round(m * 1e3, 1), "Mrays_s": round(rs.shape[0] / m / 1e6, 3)}
    return {
        "value": round(mesh.n / med / 1e6, 3), "unit": "Mtri/s", "cores": int(threads), "kind": "port",
        "sample": f"OpenMP oracle LBVH build of the same {mesh.n:,}-triangle mesh, median of 5 after 1 warmup "
                  f"({med * 1e3:.1f} ms each), OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}",
        "cpu": cpu_model(),
        "companions_omp": extra,
        "serial": serial,
    }, parity


def calib_copy(bytes_=256 << 20, reps=3):
    import ctypes as C
    import bvh_amd as B
    from bvh_amd import _native as N
    ctx = B.Context(0)
    src, dst = C.c_void_p(), C.c_void_p()
    lib = N.hip()
    lib.bvh_device_alloc(ctx.h, bytes_, C.byref(src))
    lib.bvh_device_alloc(ctx.h, bytes_, C.byref(dst))
    ms = C.c_float(0)
    for _ in range(reps):
        st = lib.bvh_debug_stream_copy(ctx.h, dst, src, bytes_, C.byref(ms))
        if st != 0:
            raise N.NativeError(st, "stream copy")
    lib.bvh_device_free(ctx.h, src)
    lib.bvh_device_free(ctx.h, dst)
    ctx.close()
    print(json.dumps({"calib_copy_bytes": bytes_, "ms": ms.value,
                      "GBps": 2 * bytes_ / (ms.value * 1e-3) / 1e9}), flush=True)


def run_sections(args):
    """Selected workloads only, for profiling (kernel traces, counters)."""
    import bvh_amd as B
    want = {x.strip() for x in args.sections.split(",") if x.strip()}
    unknown = want - {"companions", "batch", "batch_eager", "large", "serve"}
    if unknown:
        raise SystemExit(f"unknown sections: {sorted(unknown)}")
    out = {}
    ctx = B.Context(0)
    if "companions" in want:
        log(f"[bench] generating {args.n:,} blob triangles (seed 7)")
        mesh = B.Mesh.blobs(7, args.n)
        dm = ctx.upload(mesh)
        tree = ctx.tree()
        out["config1_companions"] = extras(ctx, dm, tree, mesh, args.steps)
        tree.free()
        dm.free()
    if "batch" in want:
        out["batched_vs_plain"] = batched_vs_plain(ctx, args.steps)
    if "batch_eager" in want:
        out["batched_eager"] = batched_eager(ctx, args.steps)
    if "large" in want:
        out["large_32m_strands"] = large_strands(ctx, max(1, args.steps // 4))
    ctx.close()
    if "serve" in want:
        out["serving"] = serving_latency()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lbvh-only", action="store_true", help="skip the LBVH+/PLOC/trace companions")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-latency measurement")
    ap.add_argument("--no-large", action="store_true", help="skip the 32M-triangle strands workload")
    ap.add_argument("--calib", action="store_true",
                    help="only run the 256 MiB stream-copy kernel (reference point for HBM counters)")
    ap.add_argument("--serve-child", default="", help=argparse.SUPPRESS)
    ap.add_argument("--sections", default="",
                    help="profiling: comma list of companions,batch,batch_eager,large,serve to run without the headline "
                         "(prints their JSON only)")
    args = ap.parse_args()
    if args.serve_child:
        serve_child(json.loads(args.serve_child))
        return
    if args.calib:
        calib_copy()
        return
    if args.sections:
        run_sections(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this engine shards by mesh/ray "
              f"batches across GPUs, which is not built here).", file=sys.stderr)
        sys.exit(2)

    import bvh_amd as B
    log(f"[bench] generating {args.n:,} blob triangles (seed 7)")
    mesh = B.Mesh.blobs(7, args.n)
    log(f"[bench] LBVH: {args.warmup} warmup + {args.steps} timed builds")
    ctx = B.Context(0)
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    wall, dev_ms, stages, _ = timed_builds(ctx, dm, tree, args.steps, args.warmup, algo=B.ALGO_LBVH)
    nodes, idx = tree.download()
    extra = extras(ctx, dm, tree, mesh, args.steps) if not args.lbvh_only else {}
    tree.free()
    dm.free()
    batched = batched_vs_plain(ctx, args.steps) if not (args.lbvh_only or args.no_batch) else None
    large = large_strands(ctx) if not (args.lbvh_only or args.no_large) else None
    ctx.close()
    served = serving_latency() if not (args.lbvh_only or args.no_serve) else None
    ms_per_step = wall / args.steps * 1e3
    value = mesh.n / (ms_per_step / 1e3) / 1e6

    # Per-stage algorithmic bandwidth; the dominant kernel by total time.
    per_stage = {}
    kernels = {}
    for name, ms in stages.items():
        ub = stage_unit_bytes(name)
        gbs = (ub * mesh.n) / (ms * 1e-3) / 1e9 if ub and ms > 0 else None
        per_stage[name] = {"ms": round(ms, 4), "bytes_per_unit": ub,
                           "GBps": round(gbs, 1) if gbs else None,
                           "frac_hbm": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
        kname = {"Bounds/Centroids": "k_tri_bounds", "Morton Codes": "k_morton",
                 "Topology+Refit": "k_hier"}.get(name, "k_onesweep" if name.startswith("Sort Pass") else name)
        k = kernels.setdefault(kname, {"ms": 0.0, "launches": 0, "unit": ub})
        k["ms"] += ms
        k["launches"] += 1
    dom_name, dom = max(kernels.items(), key=lambda kv: kv[1]["ms"])
    avg_launch_ms = dom["ms"] / dom["launches"]
    achieved = dom["unit"] * mesh.n / (avg_launch_ms * 1e-3) / 1e9
    traffic = pmc_traffic(dom_name, mesh.n)
    # k_hier is the hierarchy stage: k_hier_level<true,false> (level 1),
    # k_hier_level<false,false> per later level (k_hier_gather between them above
    # 8M leaves) and k_hier_climb, one stage per build.
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_name,
                "avg_launch_ms": round(avg_launch_ms, 4), "bytes_per_launch": dom["unit"] * mesh.n,
                "traffic_source": TRAFFIC_FILE_REL if traffic else None}
    if dom_name == "k_hier":
        # The same launch priced at the bytes this engine's level hierarchy
        # needs (140 B per triangle).
        a140 = DESIGN_HIER_BYTES * mesh.n / (avg_launch_ms * 1e-3) / 1e9
        roofline["achieved_at_140B_per_tri"] = round(a140, 1)
        roofline["frac_at_140B_per_tri"] = round(a140 / HBM_PEAK_GBS, 4)

    cpu = None
    parity = None
    if not args.no_cpu:
        log("[bench] CPU baseline (OpenMP oracle)")
        cpu, parity = cpu_baseline(mesh, nodes, idx)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mtri/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"LBVH build of {mesh.n:,} fp32 blob triangles (64 Gaussian blobs sigma 0.02 + 5% "
                               f"large triangles, seed 7), mesh resident in HBM", "triangles": mesh.n,
                   "algo": "LBVH", "parallelism": "single GPU"},
        "device_ms_per_build": round(dev_ms, 4),
        "device_Mtri_s": round(mesh.n / dev_ms / 1e3, 2),
        "stages": per_stage,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity_vs_oracle": parity,
        "config1_companions": extra,
        "batched_vs_plain": batched,
        "serving": served,
        "large_32m_strands": large,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
