"""Batched segmented builds on the GPU: every tree of a batch must be
bit-identical to the standalone build of its mesh (which the other GPU tests
pin to the oracle), and small meshes are also checked against the oracle
directly.  Covers graph replays, uneven and tiny meshes, meshes that need
global PLOC iterations, tuning invariance and input errors."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _meshes(seed, k, lo, hi):
    counts = B.batch_counts(seed, k, lo, hi)
    return [B.Mesh.batch_mesh(seed, i, int(c)) for i, c in enumerate(counts)]


def _standalone(ctx, dm, algo, radius):
    t = ctx.tree()
    try:
        ctx.build(dm, t, algo=algo, radius=radius)
        return t.download()
    finally:
        t.free()


def _run_batch(ctx, meshes, algo, radius=16, tune=False, runs=2):
    dms = [ctx.upload(m) for m in meshes]
    trees = [ctx.tree() for _ in meshes]
    batch = ctx.batch(dms, trees, algo=algo, radius=radius, tune=tune)
    try:
        results = []
        for _ in range(runs):
            batch.run()
            results.append([t.download() for t in trees])
        ref = [_standalone(ctx, dm, algo, radius) for dm in dms]
        info = batch.info()
    finally:
        batch.free()
        for t in trees:
            t.free()
        for d in dms:
            d.free()
    return results, ref, info


def _assert_identical(results, ref):
    for r, res in enumerate(results):
        for s, ((nodes, idx), (rn, ri)) in enumerate(zip(res, ref)):
            np.testing.assert_array_equal(idx, ri, err_msg=f"run {r} mesh {s} indices")
            assert nodes.tobytes() == rn.tobytes(), f"run {r} mesh {s}: {O.nodes_equal(nodes, rn)}"


@pytest.mark.parametrize("radius", [16, 100])
def test_batch_ploc_large_segments(ctx, radius):
    """Four 300K-triangle meshes: the batch's first iterations use the
    large-m (segmented) neighbour search, each mesh alone the small-m one."""
    meshes = [B.Mesh.batch_mesh(31, i, 300_000) for i in range(4)]
    results, ref, info = _run_batch(ctx, meshes, B.ALGO_PLOC, radius=radius, runs=1)
    _assert_identical(results, ref)


@pytest.mark.parametrize("algo", [B.ALGO_LBVH, B.ALGO_PLOC])
def test_batch_uneven_meshes(ctx, algo):
    meshes = _meshes(11, 24, 1000, 30000)
    results, ref, info = _run_batch(ctx, meshes, algo)
    _assert_identical(results, ref)
    assert info["meshes"] == 24 and info["triangles"] == sum(m.n for m in meshes)
    assert info["graph_nodes"] > 0


@pytest.mark.parametrize("algo", [B.ALGO_LBVH, B.ALGO_PLOC])
def test_batch_tiny_meshes_vs_oracle(ctx, algo):
    sizes = [1, 2, 3, 5, 64, 65, 257, 1, 2049, 4097, 17, 2]
    meshes = [B.Mesh.uniform(500 + i, n) for i, n in enumerate(sizes)]
    results, ref, _ = _run_batch(ctx, meshes, algo, radius=3)
    _assert_identical(results, ref)
    for m, (nodes, idx) in zip(meshes, results[-1]):
        if algo == B.ALGO_LBVH:
            onodes, oidx, _ = O.build_lbvh(m.soa)
        else:
            st, onodes, oidx, _ = O.build_ploc(m.soa, 3)
            assert st == 0
        np.testing.assert_array_equal(idx, oidx)
        assert O.nodes_equal(nodes, onodes) is None


def test_batch_ploc_global_iterations(ctx):
    # Two meshes above the LDS-tail size force captured global iterations
    # that run over all meshes at once.
    meshes = [B.Mesh.blobs(21, 150_000), B.Mesh.uniform(22, 3000), B.Mesh.uniform(23, 90_000)]
    results, ref, info = _run_batch(ctx, meshes, B.ALGO_PLOC, radius=16)
    _assert_identical(results, ref)
    assert info["ploc_iterations"] > 0
    assert info["ploc_captured_iterations"] == info["ploc_iterations"]
    assert info["eager_continuations"] == 0


def test_batch_single_mesh(ctx):
    meshes = [B.Mesh.uniform(31, 70_000)]
    for algo in (B.ALGO_LBVH, B.ALGO_PLOC):
        results, ref, _ = _run_batch(ctx, meshes, algo)
        _assert_identical(results, ref)


@pytest.mark.parametrize("algo,knobs", [(B.ALGO_LBVH, 3), (B.ALGO_PLOC, 5)])
def test_batch_tuning_is_result_invariant(ctx, algo, knobs):
    """Coordinate descent over digit width, sort tile, front-end chunk (and for
    PLOC the search and merge grids): one baseline plus one alternative per
    knob, every candidate's trees equal (checked inside by digest), the chosen
    configuration's trees equal to standalone builds."""
    meshes = _meshes(5, 16, 1000, 20000)
    results, ref, info = _run_batch(ctx, meshes, algo, tune=True, runs=1)
    _assert_identical(results, ref)
    assert info["tune_count"] == 1 + knobs
    tried = info["tune"]
    assert {t["radix_bits"] for t in tried} == {8, 10}
    assert {t["sort_block"] for t in tried} == {256, 512}
    assert {t["front_chunk"] for t in tried} == {4096, 8192}
    if algo == B.ALGO_PLOC:
        assert {t["nn_grid"] for t in tried} == {1024, 2048}
        assert {t["merge_grid"] for t in tried} == {512, 1024}
    best = min(tried, key=lambda t: t["ms"])  # the first fastest, as the tuner keeps it
    assert (info["radix_bits"], info["sort_block"], info["front_chunk"]) == \
        (best["radix_bits"], best["sort_block"], best["front_chunk"])
    if algo == B.ALGO_PLOC:
        assert (info["ploc_nn_grid"], info["ploc_merge_grid"]) == (best["nn_grid"], best["merge_grid"])


def test_batch_rejects_nonfinite(ctx):
    good = B.Mesh.uniform(41, 500)
    soa = B.Mesh.uniform(42, 500).soa.copy()
    soa[4, 7] = np.inf
    bad = B.Mesh(soa)
    dms = [ctx.upload(good), ctx.upload(bad)]
    trees = [ctx.tree(), ctx.tree()]
    try:
        with pytest.raises(B.NativeError) as ei:
            b = ctx.batch(dms, trees, algo=B.ALGO_LBVH)
            try:
                b.run()
            finally:
                b.free()
        assert ei.value.status == 4 and "mesh 1" in str(ei.value)
    finally:
        for t in trees:
            t.free()
        for d in dms:
            d.free()


def test_batch_rejects_lbvh_plus(ctx):
    dm = ctx.upload(B.Mesh.uniform(43, 100))
    t = ctx.tree()
    try:
        with pytest.raises(B.NativeError) as ei:
            ctx.batch([dm], [t], algo=B.ALGO_LBVH_PLUS)
        assert ei.value.status == 9
    finally:
        t.free()
        dm.free()


@pytest.mark.parametrize("algo", [B.ALGO_LBVH, B.ALGO_PLOC])
def test_build_batched_eager(ctx, algo):
    # Eager segmented build (no graph): same trees as standalone builds.
    meshes = _meshes(17, 12, 500, 60000)
    dms = [ctx.upload(m) for m in meshes]
    trees = [ctx.tree() for _ in meshes]
    try:
        t = ctx.build_batched(dms, trees, algo=algo, radius=16)
        assert t["total_ms"] > 0
        got = [tr.download() for tr in trees]
        ref = [_standalone(ctx, dm, algo, 16) for dm in dms]
    finally:
        for tr in trees:
            tr.free()
        for d in dms:
            d.free()
    _assert_identical([got], ref)


@pytest.mark.parametrize("algo", [B.ALGO_LBVH, B.ALGO_PLOC])
def test_build_batched_eager_wide_sort(ctx, algo):
    # Three 1M-triangle meshes eagerly: 3M segmented keys take the wide
    # (16384-key) sort tile; each tree equals its standalone build.
    meshes = [B.Mesh.batch_mesh(41, i, 1_000_000) for i in range(3)]
    dms = [ctx.upload(m) for m in meshes]
    trees = [ctx.tree() for _ in meshes]
    try:
        ctx.build_batched(dms, trees, algo=algo, radius=16)
        got = [tr.download() for tr in trees]
        ref = [_standalone(ctx, dm, algo, 16) for dm in dms]
    finally:
        for tr in trees:
            tr.free()
        for d in dms:
            d.free()
    _assert_identical([got], ref)


def test_build_batched_eager_large_lbvh(ctx):
    # Nine 1M-triangle meshes (8,793 leaf chunks): the hierarchy compacts the
    # leaf level in a separate pass above 8,192 chunks, while each mesh alone
    # takes the folded level 1; both must give the same trees.
    meshes = [B.Mesh.batch_mesh(43, i, 1_000_000) for i in range(9)]
    dms = [ctx.upload(m) for m in meshes]
    trees = [ctx.tree() for _ in meshes]
    try:
        ctx.build_batched(dms, trees, algo=B.ALGO_LBVH)
        got = [tr.download() for tr in trees]
        ref = [_standalone(ctx, dm, B.ALGO_LBVH, 16) for dm in dms]
    finally:
        for tr in trees:
            tr.free()
        for d in dms:
            d.free()
    _assert_identical([got], ref)


@pytest.mark.parametrize("algo", [B.ALGO_LBVH, B.ALGO_PLOC])
def test_batch_many_tiny_meshes(ctx, algo):
    # 3000 meshes of 1-40 triangles: 12 mesh-id bits in the keys, thousands of
    # one-chunk segments, every mesh finished in the per-mesh LDS tail (PLOC).
    rng = np.random.default_rng(8)
    meshes = [B.Mesh.uniform(9000 + i, int(n)) for i, n in enumerate(rng.integers(1, 41, 3000))]
    results, ref, info = _run_batch(ctx, meshes, algo, radius=4, runs=1)
    _assert_identical(results, ref)
    assert info["meshes"] == 3000 and info["key_bits"] == 30 + 12   # sorted bits: morton + mesh id
