"""Nearest-first traversal (BVH_TRACE_NEAREST) on the GPU against the oracle's
nearest-first restatement: hit ids, t bits and the three per-ray counters
identical; hits equal to the all-triangle brute force (least (t, index)) on
small meshes; the context's order switches back to the reference's."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O
from test_oracle_nearest import RAY_DOWN_Z, mesh_of, tree_of, tri_at_z

pytestmark = pytest.mark.gpu


@pytest.fixture
def nearest(ctx):
    ctx.set_trace_order(B.TRACE_NEAREST)
    try:
        yield ctx
    finally:
        ctx.set_trace_order(B.TRACE_REFERENCE)


def _nearest_both(ctx, mesh, algo, rays, brute=False, **kw):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=algo, **kw)
        nodes, _ = tree.download()
        t, prim, cnt, _ = ctx.trace(tree, dm, rays, counters=True)
    finally:
        tree.free()
        dm.free()
    ot, op, oc = O.trace(nodes, mesh.soa, rays, nearest=True)
    np.testing.assert_array_equal(prim, op)
    np.testing.assert_array_equal(t.view(np.uint32), ot.view(np.uint32))
    np.testing.assert_array_equal(cnt, oc)
    if brute:
        bt, bp = O.trace_brute(mesh.soa, rays)
        np.testing.assert_array_equal(prim, bp)
        np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
    return prim, cnt


@pytest.mark.parametrize("algo,kw", [(B.ALGO_LBVH, {}), (B.ALGO_PLOC, {"radius": 25}),
                                     (B.ALGO_LBVH_PLUS, {"rounds": 2})])
def test_nearest_small_meshes_match_oracle_and_brute_force(nearest, algo, kw):
    for n in (1, 2, 100, 5000):
        m = B.Mesh.uniform(n, n)
        rays = B.rays_sphere_to_box(n, m.bounds(), 20000)
        _nearest_both(nearest, m, algo, rays, brute=True, **kw)


@pytest.mark.parametrize("algo,kw", [(B.ALGO_LBVH, {}), (B.ALGO_PLOC, {"radius": 25}),
                                     (B.ALGO_LBVH_PLUS, {"rounds": 3})])
def test_nearest_camera_blobs_1m(nearest, algo, kw):
    m = B.Mesh.blobs(7, 1_000_000)
    rays = B.rays_sphere_camera(m.bounds(), 512, 384)
    prim, cnt = _nearest_both(nearest, m, algo, rays, **kw)
    assert (prim >= 0).mean() > 0.05


@pytest.mark.parametrize("algo,kw", [(B.ALGO_PLOC, {"radius": 100}), (B.ALGO_LBVH_PLUS, {"rounds": 3})])
def test_nearest_incoherent_strands(nearest, algo, kw):
    m = B.Mesh.strands(99, 2000, 160)
    rays = B.rays_uniform_cube(99, 100_000)
    _nearest_both(nearest, m, algo, rays, **kw)


def _identical(n):
    return B.Mesh(np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32),
                          (1, n)))


def test_nearest_deep_chain(nearest):
    # A PLOC chain about 200 deep over identical triangles: every hit ties in
    # t, so the least index wins, through the spill slots and the deep path.
    m = _identical(200)
    rays = B.rays_sphere_to_box(5, m.bounds(), 2000)
    prim, _ = _nearest_both(nearest, m, B.ALGO_PLOC, rays, brute=True, radius=2)
    assert (prim >= 0).mean() > 0.1


def test_nearest_too_deep_reported(nearest):
    m = _identical(600)
    dm = nearest.upload(m)
    tree = nearest.tree()
    try:
        nearest.build(dm, tree, algo=B.ALGO_PLOC, radius=2)
        rays = B.rays_sphere_to_box(6, m.bounds(), 500)
        with pytest.raises(B.NativeError) as ei:
            nearest.trace(tree, dm, rays)
        assert ei.value.status == 8
    finally:
        tree.free()
        dm.free()


def test_order_is_per_context_and_switches_back(ctx):
    # Hand-built tree where the two orders pick different triangles of an
    # equal-t pair (the reference takes the first met, the nearest order the
    # smaller index), traced through bvh_tree_from_nodes.
    far = [tri_at_z(0.0, x0=50.0 + 3 * k) for k in range(6)]
    far[2] = tri_at_z(1.0)
    far[5] = tri_at_z(1.0)
    m = mesh_of(far)
    nodes = tree_of((((5, 0), (1, 3)), (2, 4)), m)
    dm = ctx.upload(m)
    tree = ctx.tree()
    other = B.Context(0)
    dm2 = other.upload(m)
    tree2 = other.tree()
    try:
        tree.upload_nodes(nodes)
        tree2.upload_nodes(nodes)
        t, p, c, _ = ctx.trace(tree, dm, RAY_DOWN_Z, counters=True)
        assert p[0] == 5 and t[0] == np.float32(9.0)
        ctx.set_trace_order(B.TRACE_NEAREST)
        t, p, _ = ctx.trace(tree, dm, RAY_DOWN_Z)
        assert p[0] == 2 and t[0] == np.float32(9.0)
        t, p, _ = other.trace(tree2, dm2, RAY_DOWN_Z)   # another context keeps the reference order
        assert p[0] == 5
        ctx.set_trace_order(B.TRACE_REFERENCE)
        t, p, c2, _ = ctx.trace(tree, dm, RAY_DOWN_Z, counters=True)
        assert p[0] == 5
        np.testing.assert_array_equal(c, c2)
        with pytest.raises(B.NativeError) as ei:
            ctx.set_trace_order(7)
        assert ei.value.status == 1
    finally:
        ctx.set_trace_order(B.TRACE_REFERENCE)
        tree.free()
        dm.free()
        tree2.free()
        dm2.free()
        other.close()


def test_nearest_async_matches_synchronous(nearest):
    m = B.Mesh.blobs(301, 200_000)
    dm = nearest.upload(m)
    tree = nearest.tree()
    rays = B.rays_sphere_to_box(302, m.bounds(), 50_000)
    n = rays.shape[0]
    bufs = [nearest.buffer(n * 24), nearest.buffer(n * 4), nearest.buffer(n * 4), nearest.buffer(n * 12)]
    try:
        bufs[0].write(rays)
        nearest.build_async(dm, tree, algo=B.ALGO_LBVH)
        nearest.trace_async(tree, dm, *bufs)
        nearest.sync()
        at, ap = bufs[1].read(np.float32, n), bufs[2].read(np.int32, n)
        ac = bufs[3].read(np.int32, 3 * n).reshape(n, 3)
        st, sp, sc, _ = nearest.trace(tree, dm, rays, counters=True)
        np.testing.assert_array_equal(ap, sp)
        np.testing.assert_array_equal(at.view(np.uint32), st.view(np.uint32))
        np.testing.assert_array_equal(ac, sc)
        assert (ap >= 0).mean() > 0.2
    finally:
        for b in bufs:
            b.free()
        tree.free()
        dm.free()
