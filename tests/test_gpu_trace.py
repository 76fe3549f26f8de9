"""Traversal on the GPU against the oracle: hit triangle ids identical, t
bit-identical (well inside the 1e-6 relative tolerance), per-ray counters
identical to the reference renderer's definitions."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _trace_both(ctx, mesh, algo, rays, **kw):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=algo, **kw)
        nodes, _ = tree.download()
        t, prim, cnt, ms = ctx.trace(tree, dm, rays, counters=True)
    finally:
        tree.free()
        dm.free()
    ot, op, oc = O.trace(nodes, mesh.soa, rays, omp=True)
    np.testing.assert_array_equal(prim, op)
    hit = op >= 0
    assert np.all(np.abs(t[hit] - ot[hit]) <= 1e-6 * np.abs(ot[hit]))
    np.testing.assert_array_equal(t.view(np.uint32), ot.view(np.uint32))
    np.testing.assert_array_equal(cnt, oc)
    return hit.mean(), ms


@pytest.mark.parametrize("algo,kw", [(B.ALGO_LBVH, {}), (B.ALGO_PLOC, {"radius": 25}),
                                     (B.ALGO_LBVH_PLUS, {"rounds": 2})])
def test_trace_small_meshes(ctx, algo, kw):
    for n in (1, 2, 100, 5000):
        m = B.Mesh.uniform(n, n)
        rays = B.rays_sphere_to_box(n, m.bounds(), 20000)
        _trace_both(ctx, m, algo, rays, **kw)


def test_trace_camera_blobs_1m(ctx):
    m = B.Mesh.blobs(7, 1_000_000)
    rays = B.rays_sphere_camera(m.bounds(), 512, 384)
    frac, _ = _trace_both(ctx, m, B.ALGO_LBVH, rays)
    assert frac > 0.05


def test_trace_incoherent_strands(ctx):
    m = B.Mesh.strands(99, 2000, 160)
    rays = B.rays_uniform_cube(99, 100_000)
    _trace_both(ctx, m, B.ALGO_PLOC, rays, radius=100)


def test_trace_external_nodes(ctx):
    """A node array produced elsewhere (here: the oracle) traced through bvh_tree_from_nodes."""
    m = B.Mesh.uniform(3, 20000)
    onodes, _, _ = O.build_lbvh(m.soa)
    dm = ctx.upload(m)
    tree = ctx.tree()
    try:
        tree.upload_nodes(onodes)
        rays = B.rays_sphere_to_box(4, m.bounds(), 10000)
        t, prim, ms = ctx.trace(tree, dm, rays)
    finally:
        tree.free()
        dm.free()
    ot, op, _ = O.trace(onodes, m.soa, rays)
    np.testing.assert_array_equal(prim, op)
    np.testing.assert_array_equal(t.view(np.uint32), ot.view(np.uint32))


def _identical(n):
    return B.Mesh(np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32),
                          (1, n)))


def test_trace_deep_chain_uses_full_stack(ctx):
    # PLOC on identical triangles merges one pair per iteration: a chain about
    # 200 deep, so rays through it need far more than the 24 LDS entries.
    m = _identical(200)
    rays = B.rays_sphere_to_box(5, m.bounds(), 2000)
    frac, _ = _trace_both(ctx, m, B.ALGO_PLOC, rays, radius=2)
    assert frac > 0.1


def test_trace_reports_too_deep(ctx):
    m = _identical(600)
    dm = ctx.upload(m)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=B.ALGO_PLOC, radius=2)
        rays = B.rays_sphere_to_box(6, m.bounds(), 500)
        with pytest.raises(B.NativeError) as ei:
            ctx.trace(tree, dm, rays)
        assert ei.value.status == 8
    finally:
        tree.free()
        dm.free()


def test_tree_save_load_and_viz_export(ctx, tmp_path):
    m = B.Mesh.blobs(12, 30000)
    dm = ctx.upload(m)
    t1, t2 = ctx.tree(), ctx.tree()
    try:
        ctx.build(dm, t1, algo=B.ALGO_PLOC, radius=16)
        nodes, idx = t1.download()
        p = str(tmp_path / "t.bvh")
        t1.save(p)
        t2.load(p)
        n2, i2 = t2.download()
        assert n2.tobytes() == nodes.tobytes()
        np.testing.assert_array_equal(i2, idx)
        rays = B.rays_sphere_to_box(12, m.bounds(), 20000)
        ta, pa, _ = ctx.trace(t1, dm, rays)
        tb, pb, _ = ctx.trace(t2, dm, rays)
        np.testing.assert_array_equal(pa, pb)
        assert ta.tobytes() == tb.tobytes()
        v = str(tmp_path / "t.viz")
        t1.export_viz(v)
        viz = np.fromfile(v, dtype=np.dtype([("mn", "<f4", (3,)), ("mx", "<f4", (3,)), ("l", "<i4"), ("r", "<i4")]))
        np.testing.assert_array_equal(viz["mn"], nodes["bmin"])
        np.testing.assert_array_equal(viz["mx"], nodes["bmax"])
        np.testing.assert_array_equal(viz["l"], nodes["left"])
        np.testing.assert_array_equal(viz["r"], nodes["right"])
    finally:
        t1.free()
        t2.free()
        dm.free()
