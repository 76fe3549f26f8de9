"""Serving loop on the GPU: every request's hits are bit-identical whether it
runs alone (plain build + trace on its own context) or inside a mix of
concurrent, coalesced requests on several slots; residency, eviction and
budget rejection."""
import numpy as np
import pytest

import bvh_amd as B
from bvh_amd import serving as S

pytestmark = pytest.mark.gpu


def _alone(ctx, mesh, rays, algo, radius):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=algo, radius=radius)
        t, prim, _ = ctx.trace(tree, dm, rays)
    finally:
        tree.free()
        dm.free()
    return t, prim


@pytest.mark.parametrize("algo", [B.ALGO_LBVH, B.ALGO_PLOC])
def test_serving_same_alone_and_in_mix(ctx, algo):
    reqs = S.make_requests(3, 18, tri_lo=2000, tri_hi=150_000, ray_lo=2000, ray_hi=40_000, follow_up=0.25)
    with S.Server(algo=algo, radius=16, slots=3, max_group=6) as srv:
        done, outs, _ = S.run_open_loop(srv, reqs, np.zeros(len(reqs)))
        stats = srv.stats()
    assert [c["status"] for c in done] == [0] * len(reqs), done
    assert stats["max_group_seen"] > 1, stats
    assert stats["completed"] == len(reqs)
    mesh_of = {q["key"]: q["mesh"] for q in reqs if q["mesh"] is not None}
    for q in reqs:
        t, prim = _alone(ctx, mesh_of[q["key"]], q["rays"], algo, 16)
        got_t, got_prim = outs[q["id"]]
        np.testing.assert_array_equal(got_prim, prim, err_msg=f"request {q['id']}")
        assert got_t.tobytes() == t.tobytes(), f"request {q['id']}"
    assert sum(c["built"] for c in done) == sum(1 for q in reqs if q["mesh"] is not None)


def test_serving_nearest_order_same_alone_and_in_mix(ctx):
    """A server configured with the nearest-first order: each request's hits
    equal its own build traced alone in that order, and differ from the
    reference order's only at equal t."""
    reqs = S.make_requests(5, 12, tri_lo=2000, tri_hi=150_000, ray_lo=2000, ray_hi=40_000, follow_up=0.25)
    with S.Server(algo=B.ALGO_LBVH, slots=3, max_group=6, trace_order=B.TRACE_NEAREST) as srv:
        done, outs, _ = S.run_open_loop(srv, reqs, np.zeros(len(reqs)))
    assert [c["status"] for c in done] == [0] * len(reqs), done
    mesh_of = {q["key"]: q["mesh"] for q in reqs if q["mesh"] is not None}
    for q in reqs:
        ctx.set_trace_order(B.TRACE_NEAREST)
        try:
            t, prim = _alone(ctx, mesh_of[q["key"]], q["rays"], B.ALGO_LBVH, 16)
        finally:
            ctx.set_trace_order(B.TRACE_REFERENCE)
        got_t, got_prim = outs[q["id"]]
        np.testing.assert_array_equal(got_prim, prim, err_msg=f"request {q['id']}")
        assert got_t.tobytes() == t.tobytes(), f"request {q['id']}"
        rt, rprim = _alone(ctx, mesh_of[q["key"]], q["rays"], B.ALGO_LBVH, 16)
        other = got_prim != rprim
        assert np.array_equal(got_t[other], rt[other]), f"request {q['id']}"


def test_serving_resident_followups_and_eviction(ctx):
    meshes = [B.Mesh.blobs(60 + i, 50_000) for i in range(4)]
    tr, rs = S.footprint(50_000, 1000, B.ALGO_LBVH)
    budget = 2 * rs + 2 * tr  # room for two resident meshes while one more builds
    with S.Server(algo=B.ALGO_LBVH, slots=1, max_group=1, budget=budget) as srv:
        for i, m in enumerate(meshes):
            rays = B.rays_sphere_to_box(i, m.bounds(), 1000)
            t, prim = srv.submit(i, m, rays, key=100 + i)
            srv.drain()
            srv.poll(0)
        st = srv.stats()
        assert st["evictions"] >= 1 and st["resident_meshes"] <= 3, st
        # the most recent mesh is still resident: a follow-up needs no upload
        m = meshes[-1]
        rays = B.rays_sphere_to_box(9, m.bounds(), 2000)
        t, prim = srv.submit(50, None, rays, key=103)
        srv.drain()
        c = srv.poll(100)
        assert c and c[0]["status"] == 0 and c[0]["built"] == 0
        t_ref, prim_ref = _alone(ctx, m, rays, B.ALGO_LBVH, 16)
        np.testing.assert_array_equal(prim, prim_ref)


def test_server_restores_the_cache_limit(ctx):
    before = ctx.cache_set_limit(3 << 30)
    try:
        with S.Server(algo=B.ALGO_LBVH, slots=1, budget=1 << 20):
            pass
        assert ctx.cache_set_limit(3 << 30) == 3 << 30   # the server's 1/8 of 1 MiB is gone again
    finally:
        ctx.cache_set_limit(before)


def test_serving_rejections(ctx):
    with S.Server(algo=B.ALGO_LBVH, slots=1, budget=1 << 20) as srv:
        m = B.Mesh.uniform(3, 100_000)
        srv.submit(1, m, B.rays_sphere_to_box(1, m.bounds(), 100), key=0)
        srv.submit(2, None, B.rays_sphere_to_box(2, m.bounds(), 100), key=12345)
        srv.drain()
        got = {c["id"]: c["status"] for c in srv.poll(100)}
        assert got == {1: 6, 2: 8}
        assert srv.stats()["rejected"] == 2


@pytest.mark.parametrize("algo,sizes,nrays", [(B.ALGO_LBVH, [300_000], 70_000), (B.ALGO_PLOC, [300_000], 20_000),
                                              (B.ALGO_LBVH_PLUS, [250_000], 5_000),
                                              (B.ALGO_LBVH, [1000, 50_000, 123_457, 7], 100_000),
                                              (B.ALGO_PLOC, [30_000, 2_000, 90_001], 1_000)])
def test_held_workspace_within_footprint(algo, sizes, nrays):
    """After real builds and a trace, a fresh context holds no more device
    memory than bvh_footprint_estimate says (what admission charges)."""
    import ctypes as C
    from bvh_amd import _native as N
    with B.Context(0) as c:
        meshes = [B.Mesh.blobs(40 + i, n) for i, n in enumerate(sizes)]
        dms = [c.upload(m) for m in meshes]
        trees = [c.tree() for _ in dms]
        try:
            if len(dms) == 1:
                c.build(dms[0], trees[0], algo=algo, radius=16, rounds=2)
            else:
                c.build_batched(dms, trees, algo=algo, radius=16)
            rays = B.rays_sphere_to_box(3, meshes[0].bounds(), nrays)
            c.trace(trees[0], dms[0], rays)
            hb, ht = C.c_uint64(), C.c_uint64()
            assert N.hip().bvh_ctx_held_bytes(c.h, C.byref(hb), C.byref(ht)) == 0
            f = N.Footprint()
            assert N.hip().bvh_footprint_estimate(algo, sum(sizes), len(sizes), nrays, C.byref(f)) == 0
            # The trace's ray sort reuses (and may grow) the build's digit
            # table, so the two parts are compared together.
            assert hb.value + ht.value <= f.build_workspace + f.trace_workspace, (hb.value, ht.value, f.build_workspace,
                                                                                 f.trace_workspace)
            assert hb.value >= 0.9 * f.build_workspace   # the estimate is not a loose bound either
        finally:
            for t in trees:
                t.free()
            for d in dms:
                d.free()
