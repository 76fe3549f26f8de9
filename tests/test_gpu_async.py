"""Asynchronous build and trace (bvh_build_async, bvh_trace_async, bvh_sync):
work queued back to back on one context gives, after bvh_sync, the same trees
(bit for bit) and the same hits as the synchronous calls, and deferred
failures are reported by bvh_sync in call order."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(B.ALGO_LBVH, {}), (B.ALGO_LBVH_PLUS, {"rounds": 3}), (B.ALGO_PLOC, {"radius": 25})]


def _sync_build(ctx, dm, algo, kw):
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=algo, timing=False, **kw)
        return tree.download()
    finally:
        tree.free()


def _device_rays(ctx, rays):
    n = rays.shape[0]
    bufs = [ctx.buffer(n * 24), ctx.buffer(n * 4), ctx.buffer(n * 4), ctx.buffer(n * 12)]
    bufs[0].write(rays)
    return bufs


def test_queued_builds_and_traces_match_synchronous(ctx):
    meshes = [B.Mesh.blobs(201, 150_000), B.Mesh.uniform(202, 90_001), B.Mesh.blobs(203, 300_000)]
    dms = [ctx.upload(m) for m in meshes]
    trees = [ctx.tree() for _ in CASES]
    rays = [B.rays_sphere_to_box(210 + i, m.bounds(), 40_000) for i, m in enumerate(meshes)]
    bufs = [_device_rays(ctx, r) for r in rays]
    try:
        # Everything is queued before anything is read back.
        for (algo, kw), dm, tree in zip(CASES, dms, trees):
            ctx.build_async(dm, tree, algo=algo, **kw)
        for dm, tree, (r, t, p, c) in zip(dms, trees, bufs):
            ctx.trace_async(tree, dm, r, t, p, c)
        ctx.sync()
        for (algo, kw), dm, tree, ray, (r, t, p, c), mesh in zip(CASES, dms, trees, rays, bufs, meshes):
            nodes, idx = tree.download()
            snodes, sidx = _sync_build(ctx, dm, algo, kw)
            assert O.nodes_equal(nodes, snodes) is None, algo
            np.testing.assert_array_equal(idx, sidx)
            n = ray.shape[0]
            at, ap = t.read(np.float32, n), p.read(np.int32, n)
            ac = c.read(np.int32, 3 * n).reshape(n, 3)
            st, sp, sc, _ = ctx.trace(tree, dm, ray, counters=True)
            np.testing.assert_array_equal(ap, sp)
            np.testing.assert_array_equal(at.view(np.uint32), st.view(np.uint32))
            np.testing.assert_array_equal(ac, sc)
            assert (ap >= 0).mean() > 0.2
    finally:
        for bl in bufs:
            for b in bl:
                b.free()
        for tree in trees:
            tree.free()
        for dm in dms:
            dm.free()


def test_rebuilds_into_one_tree_run_in_call_order(ctx):
    a, b = B.Mesh.uniform(221, 60_000), B.Mesh.blobs(222, 80_000)
    da, db = ctx.upload(a), ctx.upload(b)
    tree = ctx.tree()
    try:
        ctx.build_async(da, tree, algo=B.ALGO_PLOC, radius=16)
        ctx.build_async(db, tree, algo=B.ALGO_LBVH)
        ctx.sync()
        nodes, idx = tree.download()
        onodes, oidx, _ = O.build_lbvh(b.soa)
        assert O.nodes_equal(nodes, onodes) is None
        np.testing.assert_array_equal(idx, oidx)
    finally:
        tree.free()
        da.free()
        db.free()


@pytest.mark.parametrize("algo,kw", CASES)
def test_nonfinite_mesh_refused_before_any_launch(ctx, algo, kw):
    """The upload finds a non-finite coordinate; synchronous and asynchronous
    builds of that mesh are refused at once, and queued work is unaffected."""
    good = B.Mesh.uniform(231, 20_000)
    bad = B.Mesh.uniform(232, 20_000)
    bad.soa[7, 1234] = np.inf
    dg, dbad = ctx.upload(good), ctx.upload(bad)
    t1, t2, t3 = ctx.tree(), ctx.tree(), ctx.tree()
    try:
        assert B.N.hip().bvh_mesh_check(dbad.h) == 4 and B.N.hip().bvh_mesh_check(dg.h) == 0
        ctx.build_async(dg, t1, algo=algo, **kw)
        with pytest.raises(B.NativeError) as e:
            ctx.build_async(dbad, t2, algo=algo, **kw)
        assert e.value.status == 4
        with pytest.raises(B.NativeError) as e:
            ctx.build(dbad, t2, algo=algo, **kw)
        assert e.value.status == 4
        ctx.build_async(dg, t3, algo=algo, **kw)
        ctx.sync()
        ref = _sync_build(ctx, dg, algo, kw)
        for t in (t1, t3):
            nodes, idx = t.download()
            assert O.nodes_equal(nodes, ref[0]) is None
            np.testing.assert_array_equal(idx, ref[1])
    finally:
        for t in (t1, t2, t3):
            t.free()
        dg.free()
        dbad.free()


def test_deferred_trace_failure_reported_once_in_order(ctx):
    """A trace whose rays need more than 256 stack entries (a 600-deep PLOC
    chain over identical triangles) fails at bvh_sync; the traces queued
    around it are complete."""
    chain = B.Mesh(np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32),
                           (1, 600)))
    good = B.Mesh.blobs(235, 50_000)
    dc, dg = ctx.upload(chain), ctx.upload(good)
    tc, tg = ctx.tree(), ctx.tree()
    rc = B.rays_sphere_to_box(236, chain.bounds(), 500)
    rg = B.rays_sphere_to_box(237, good.bounds(), 20_000)
    bc, bg1, bg2 = _device_rays(ctx, rc), _device_rays(ctx, rg), _device_rays(ctx, rg)
    try:
        ctx.build(dc, tc, algo=B.ALGO_PLOC, radius=2)
        ctx.build_async(dg, tg, algo=B.ALGO_LBVH)
        ctx.trace_async(tg, dg, *bg1)
        ctx.trace_async(tc, dc, *bc)
        ctx.trace_async(tg, dg, *bg2)
        with pytest.raises(B.NativeError) as e:
            ctx.sync()
        assert e.value.status == 8
        ctx.sync()
        st, sp, sc, _ = ctx.trace(tg, dg, rg, counters=True)
        n = rg.shape[0]
        for _, t, p, c in (bg1, bg2):
            np.testing.assert_array_equal(p.read(np.int32, n), sp)
            np.testing.assert_array_equal(t.read(np.float32, n).view(np.uint32), st.view(np.uint32))
            np.testing.assert_array_equal(c.read(np.int32, 3 * n).reshape(n, 3), sc)
    finally:
        for bl in (bc, bg1, bg2):
            for b in bl:
                b.free()
        tc.free()
        tg.free()
        dc.free()
        dg.free()


def test_many_pending_traces(ctx):
    """More queued calls than one pinned block of status slots holds (256)."""
    m = B.Mesh.blobs(241, 30_000)
    dm = ctx.upload(m)
    tree = ctx.tree()
    rays = B.rays_sphere_to_box(242, m.bounds(), 2_000)
    r, t, p, c = _device_rays(ctx, rays)
    try:
        ctx.build_async(dm, tree, algo=B.ALGO_LBVH)
        for _ in range(700):
            ctx.trace_async(tree, dm, r, t, p, c)
        ctx.sync()
        st, sp, _ = ctx.trace(tree, dm, rays)
        np.testing.assert_array_equal(p.read(np.int32, rays.shape[0]), sp)
        np.testing.assert_array_equal(t.read(np.float32, rays.shape[0]).view(np.uint32), st.view(np.uint32))
    finally:
        for b in (r, t, p, c):
            b.free()
        tree.free()
        dm.free()


def test_async_argument_errors_return_at_once(ctx):
    m = B.Mesh.uniform(251, 100)
    dm = ctx.upload(m)
    tree = ctx.tree()
    try:
        with pytest.raises(B.NativeError) as e:
            ctx.build_async(dm, tree, algo=B.ALGO_PLOC, radius=0)
        assert e.value.status == 1
        with pytest.raises(B.NativeError) as e:
            ctx.build_async(dm, tree, algo=7)
        assert e.value.status == 9
        ctx.sync()
    finally:
        tree.free()
        dm.free()
