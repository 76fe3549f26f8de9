"""Parity at the workloads' own sizes.

* 4,000,000 blob triangles (seed 7): LBVH+ (7-leaf treelets, 3 rounds) and
  PLOC (r = 25) bit-exact against the oracle; 1024x768 bounding-sphere camera
  rays through the LBVH, LBVH+ and PLOC trees against the oracle traversal
  (ids identical, t bit-identical; the stated tolerance is 1e-6 relative).
* Random-walk strands: LBVH+ 3 rounds and PLOC r = 100 bit-exact against the
  oracle at 1,000,000 triangles.
* 32,000,000 strand triangles (200,000 strands x 160 steps, seed 99): PLOC
  r = 100 and LBVH+ 3 rounds checked through size-independent properties
  (full binary tree reached from the root, every primitive in exactly one
  leaf, leaf boxes = triangle boxes, every internal box = the select union of
  its children, SAH non-increasing over treelet rounds), and a 65,536-ray
  sample of the 4,194,304 incoherent rays traced by the oracle through the
  same trees.
* 512 meshes of 1,000-200,000 triangles (seed 11), LBVH and PLOC r = 16 in
  one tuned captured launch sequence: every tree bit-identical to its own
  standalone build, the smallest ones also to the oracle.
* The nearest-first trace order on the same 4M and 32M cases against the
  oracle's nearest-first restatement (ids and t bits), and against the
  reference order: a different triangle only at an equal t.
* Serving requests drawn from the serving workload's ranges (10,000 to
  2,000,000 triangles, 65,536 to 1,048,576 rays, seed 23): each result in the
  mix equal to the request alone.

Oracle references: lbvh_plus_builder.cu:341-366,483-495 (rounds),
ploc_builder.cu:489-548 (iteration loop), render_engine.cu:289-356 (traversal).
"""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O
from tree_props import check_tree_properties, sah64

pytestmark = pytest.mark.gpu

T_REL = 1e-6   # stated tolerance for hit distances (the kernels give identical bits)


def _build(ctx, dm, algo, **kw):
    tree = ctx.tree()
    ctx.build(dm, tree, algo=algo, **kw)
    return tree


def _compare_trace(ctx, tree, dm, nodes, mesh, rays, sample=None):
    t, prim, _ = ctx.trace(tree, dm, rays)
    if sample is not None:
        rays, t, prim = rays[sample], t[sample], prim[sample]
    ot, op, _ = O.trace(nodes, mesh.soa, rays, omp=True)
    np.testing.assert_array_equal(prim, op)
    hit = op >= 0
    assert np.all(np.abs(t[hit] - ot[hit]) <= T_REL * np.abs(ot[hit]))
    np.testing.assert_array_equal(t.view(np.uint32), ot.view(np.uint32))
    return hit.mean()


def _compare_nearest(ctx, tree, dm, nodes, mesh, rays, sample=None):
    """The nearest-first order (BVH_TRACE_NEAREST) against the oracle's, and
    against the reference order's hits: the same triangle, or one at the
    same t (an equal-t tie, which the two orders break differently)."""
    rt, rp, _ = ctx.trace(tree, dm, rays)
    ctx.set_trace_order(B.TRACE_NEAREST)
    try:
        t, prim, _ = ctx.trace(tree, dm, rays)
    finally:
        ctx.set_trace_order(B.TRACE_REFERENCE)
    if sample is not None:
        rays, t, prim, rt, rp = rays[sample], t[sample], prim[sample], rt[sample], rp[sample]
    ot, op, _ = O.trace(nodes, mesh.soa, rays, nearest=True)
    np.testing.assert_array_equal(prim, op)
    np.testing.assert_array_equal(t.view(np.uint32), ot.view(np.uint32))
    other = prim != rp
    np.testing.assert_array_equal(t[other].view(np.uint32), rt[other].view(np.uint32))
    assert other.mean() < 1e-3
    return other.sum()


# ---------------------------------------------------------------- 4M blobs --

@pytest.fixture(scope="module")
def blobs4m(ctx):
    mesh = B.Mesh.blobs(7, 4_000_000)
    dm = ctx.upload(mesh)
    yield mesh, dm
    dm.free()


def test_blobs_4m_lbvh_plus_three_rounds(ctx, blobs4m):
    mesh, dm = blobs4m
    tree = _build(ctx, dm, B.ALGO_LBVH_PLUS, rounds=3)
    try:
        nodes, idx = tree.download()
        onodes, oidx, _ = O.build_lbvh_plus(mesh.soa, 3)
        np.testing.assert_array_equal(idx, oidx)
        assert O.nodes_equal(nodes, onodes) is None
        s_gpu, s_ref = O.sah(nodes), O.sah(onodes)
        assert abs(s_gpu - s_ref) <= 1e-5 * abs(s_ref)
        rays = B.rays_sphere_camera(mesh.bounds(), 1024, 768)
        assert _compare_trace(ctx, tree, dm, onodes, mesh, rays) > 0.05
    finally:
        tree.free()


def test_blobs_4m_ploc_r25(ctx, blobs4m):
    mesh, dm = blobs4m
    tree = _build(ctx, dm, B.ALGO_PLOC, radius=25)
    try:
        nodes, idx = tree.download()
        st, onodes, oidx, _ = O.build_ploc(mesh.soa, 25)
        assert st == 0
        np.testing.assert_array_equal(idx, oidx)
        assert O.nodes_equal(nodes, onodes) is None
        rays = B.rays_sphere_camera(mesh.bounds(), 1024, 768)
        assert _compare_trace(ctx, tree, dm, onodes, mesh, rays) > 0.05
        _compare_nearest(ctx, tree, dm, onodes, mesh, rays)
    finally:
        tree.free()


def test_blobs_4m_camera_rays_lbvh(ctx, blobs4m):
    mesh, dm = blobs4m
    tree = _build(ctx, dm, B.ALGO_LBVH)
    try:
        onodes, oidx, _ = O.build_lbvh(mesh.soa)
        nodes, idx = tree.download()
        assert O.nodes_equal(nodes, onodes) is None
        rays = B.rays_sphere_camera(mesh.bounds(), 1024, 768)
        assert _compare_trace(ctx, tree, dm, onodes, mesh, rays) > 0.05
        _compare_nearest(ctx, tree, dm, onodes, mesh, rays)
    finally:
        tree.free()


# ------------------------------------------------------------ strands 1M --

@pytest.mark.parametrize("algo,kw", [(B.ALGO_LBVH_PLUS, {"rounds": 3}), (B.ALGO_PLOC, {"radius": 100})])
def test_strands_1m_against_oracle(ctx, algo, kw):
    mesh = B.Mesh.strands(99, 6250, 160)
    dm = ctx.upload(mesh)
    tree = _build(ctx, dm, algo, **kw)
    try:
        nodes, idx = tree.download()
        if algo == B.ALGO_PLOC:
            st, onodes, oidx, _ = O.build_ploc(mesh.soa, kw["radius"])
            assert st == 0
        else:
            onodes, oidx, _ = O.build_lbvh_plus(mesh.soa, kw["rounds"])
        np.testing.assert_array_equal(idx, oidx)
        assert O.nodes_equal(nodes, onodes) is None
        rays = B.rays_uniform_cube(99, 65_536)
        _compare_trace(ctx, tree, dm, onodes, mesh, rays)
    finally:
        tree.free()
        dm.free()


# ----------------------------------------------------------- strands 32M --

@pytest.fixture(scope="module")
def strands32m(ctx):
    print("[configs] generating 32,000,000 strand triangles (seed 99)", flush=True)
    mesh = B.Mesh.strands(99, 200_000, 160)
    dm = ctx.upload(mesh)
    tb = O.tri_boxes(mesh.soa)
    rays = B.rays_uniform_cube(99, 4_194_304)
    sample = np.arange(0, rays.shape[0], 64)
    yield mesh, dm, tb, rays, sample
    dm.free()


def test_strands_32m_ploc_r100(ctx, strands32m):
    mesh, dm, tb, rays, sample = strands32m
    tree = _build(ctx, dm, B.ALGO_PLOC, radius=100)
    try:
        nodes, idx = tree.download()
        check_tree_properties(nodes, idx, tb)
        _compare_trace(ctx, tree, dm, nodes, mesh, rays, sample=sample)
        _compare_nearest(ctx, tree, dm, nodes, mesh, rays, sample=sample)
    finally:
        tree.free()


def test_strands_32m_lbvh_plus_three_rounds(ctx, strands32m):
    mesh, dm, tb, rays, sample = strands32m
    costs = []
    for rounds in (0, 1, 2, 3):
        tree = _build(ctx, dm, B.ALGO_LBVH_PLUS, rounds=rounds)
        try:
            nodes, idx = tree.download()
            if rounds in (0, 3):            # ~20 s each at 32M on the host
                check_tree_properties(nodes, idx, tb)
            costs.append(sah64(nodes))
            if rounds == 3:
                _compare_trace(ctx, tree, dm, nodes, mesh, rays, sample=sample)
                _compare_nearest(ctx, tree, dm, nodes, mesh, rays, sample=sample)
        finally:
            tree.free()
        del nodes
    print("[configs] 32M LBVH+ SAH by rounds:", costs, flush=True)
    for a, b in zip(costs, costs[1:]):
        assert b <= a * (1 + 1e-9)
    assert costs[-1] < costs[0]


# ------------------------------------------------------- 512-mesh batch --

@pytest.fixture(scope="module")
def batch512(ctx):
    """The batched workload: 512 meshes, log-uniform 1,000-200,000 triangles
    (seed 11), each a uniform-cube mesh with its own seed, scale and offset."""
    counts = B.batch_counts(11, 512, 1000, 200_000)
    print(f"[configs] generating 512 meshes, {int(counts.sum()):,} triangles", flush=True)
    meshes = [B.Mesh.batch_mesh(11, i, int(c)) for i, c in enumerate(counts)]
    dms = [ctx.upload(m) for m in meshes]
    yield meshes, dms
    for dm in dms:
        dm.free()


@pytest.mark.parametrize("algo,radius", [(B.ALGO_LBVH, 25), (B.ALGO_PLOC, 16)])
def test_batch_512_meshes_one_tuned_graph(ctx, batch512, algo, radius):
    """One tuned, captured launch sequence over all 512 meshes: every tree is
    bit-identical to its own standalone build, and the smallest meshes'
    trees equal the oracle's."""
    meshes, dms = batch512
    trees = [ctx.tree() for _ in dms]
    one = ctx.tree()
    try:
        batch = ctx.batch(dms, trees, algo=algo, radius=radius, tune=True)
        try:
            batch.run()
            info = batch.info()
        finally:
            batch.free()
        assert info["meshes"] == 512 and info["tune_count"] >= 4
        for i, (dm, tree) in enumerate(zip(dms, trees)):
            nodes, idx = tree.download()
            ctx.build(dm, one, algo=algo, radius=radius, timing=False)
            snodes, sidx = one.download()
            assert nodes.tobytes() == snodes.tobytes(), f"mesh {i}"
            np.testing.assert_array_equal(idx, sidx, err_msg=f"mesh {i}")
        small = np.argsort([m.n for m in meshes])[:6]
        for i in small:
            nodes, idx = trees[i].download()
            if algo == B.ALGO_PLOC:
                st, onodes, oidx, _ = O.build_ploc(meshes[i].soa, radius)
                assert st == 0
            else:
                onodes, oidx, _ = O.build_lbvh(meshes[i].soa)
            assert O.nodes_equal(nodes, onodes) is None, f"mesh {i}"
            np.testing.assert_array_equal(idx, oidx)
    finally:
        one.free()
        for t in trees:
            t.free()


# ------------------------------------------------- serving request ranges --

def test_serving_request_ranges_same_alone_and_in_mix(ctx):
    """Requests drawn from the serving workload's ranges (blob meshes of
    10,000-2,000,000 triangles, 65,536-1,048,576 rays, 20% follow-up traces,
    seed 23), all submitted at once to four slots: each result equals a
    plain build + trace of that request alone."""
    from bvh_amd import serving as S
    reqs = S.make_requests(23, 40, follow_up=0.2)
    with S.Server(algo=B.ALGO_LBVH, slots=4, max_group=32) as srv:
        done, outs, _ = S.run_open_loop(srv, reqs, np.zeros(len(reqs)))
        stats = srv.stats()
    assert [c["status"] for c in done] == [0] * len(reqs), done
    assert stats["max_group_seen"] > 1, stats
    mesh_of = {q["key"]: q["mesh"] for q in reqs if q["mesh"] is not None}
    for q in reqs:
        m = mesh_of[q["key"]]
        dm = ctx.upload(m)
        tree = ctx.tree()
        try:
            ctx.build(dm, tree, algo=B.ALGO_LBVH, timing=False)
            t, prim, _ = ctx.trace(tree, dm, q["rays"])
        finally:
            tree.free()
            dm.free()
        got_t, got_prim = outs[q["id"]]
        np.testing.assert_array_equal(got_prim, prim, err_msg=f"request {q['id']}")
        assert got_t.tobytes() == t.tobytes(), f"request {q['id']}"
