"""Released meshes and trees (bvh_mesh_release, bvh_tree_release): their
device blocks go to the per-device cache, later uploads and trees of a
similar size reuse them, and what is built in reused memory is bit-identical
to the oracle and to a build in fresh memory.  Freed blocks return to the
device pool, whose unused memory bvh_device_memory counts as free."""
import ctypes as C

import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def roomy_cache(ctx):
    """A cache limit these tests fit in, whatever ran before; restored after."""
    prev = ctx.cache_set_limit(4 << 30)
    yield
    ctx.cache_trim()
    ctx.cache_set_limit(prev)


def _build_trace(ctx, mesh, rays, algo=B.ALGO_LBVH, **kw):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    ctx.build(dm, tree, algo=algo, **kw)
    nodes, idx = tree.download()
    t, prim, cnt, _ = ctx.trace(tree, dm, rays, counters=True)
    return dm, tree, (nodes, idx, t, prim, cnt)


def test_released_blocks_are_reused_bit_identically(ctx):
    ctx.cache_trim()
    assert ctx.cache_info() == (0, 0)
    a = B.Mesh.blobs(401, 200_000)
    ra = B.rays_sphere_to_box(401, a.bounds(), 30_000)
    dm, tree, _ = _build_trace(ctx, a, ra)
    tree.release()
    dm.release()
    cached, blocks = ctx.cache_info()
    assert blocks == 5 and cached > 200_000 * (36 + 48 + 2 * 48)   # mesh arrays, AoS copy, nodes, indices, records
    # A slightly smaller mesh takes the released blocks.
    b = B.Mesh.blobs(402, 195_000)
    rb = B.rays_sphere_to_box(402, b.bounds(), 30_000)
    dm2, tree2, got = _build_trace(ctx, b, rb)
    assert ctx.cache_info()[1] == 0, ctx.cache_info()
    # The new objects report the released blocks' sizes.
    mb, tb = C.c_uint64(), C.c_uint64()
    assert ctx._lib.bvh_mesh_bytes(dm2.h, C.byref(mb)) == 0 and ctx._lib.bvh_tree_bytes(tree2.h, C.byref(tb)) == 0
    assert mb.value + tb.value == cached
    # The same mesh built and traced in fresh memory on another context.
    other = B.Context(0)
    try:
        dm3, tree3, want = _build_trace(other, b, rb)
        tree3.free()
        dm3.free()
    finally:
        other.close()
    assert O.nodes_equal(got[0], want[0]) is None
    np.testing.assert_array_equal(got[1], want[1])
    np.testing.assert_array_equal(got[3], want[3])
    assert got[2].tobytes() == want[2].tobytes()
    np.testing.assert_array_equal(got[4], want[4])
    onodes, oidx, _ = O.build_lbvh(b.soa)
    assert O.nodes_equal(got[0], onodes) is None
    np.testing.assert_array_equal(got[1], oidx)
    tree2.free()
    dm2.free()


def test_much_smaller_requests_do_not_take_large_blocks(ctx):
    ctx.cache_trim()
    big = B.Mesh.uniform(403, 400_000)
    dm = ctx.upload(big)
    dm.release()
    before = ctx.cache_info()
    assert before[1] == 2
    small = B.Mesh.uniform(404, 20_000)
    dm2 = ctx.upload(small)          # 20x smaller: fresh blocks, the cache is untouched
    assert ctx.cache_info() == before
    dm2.release()
    assert ctx.cache_info()[1] == 4
    ctx.cache_trim()
    assert ctx.cache_info() == (0, 0)


def test_ploc_tree_in_reused_memory_matches_oracle(ctx):
    ctx.cache_trim()
    m = B.Mesh.blobs(405, 120_000)
    for _ in range(2):  # the second build runs in the first one's released blocks
        dm = ctx.upload(m)
        tree = ctx.tree()
        ctx.build(dm, tree, algo=B.ALGO_PLOC, radius=25)
        nodes, idx = tree.download()
        tree.release()
        dm.release()
    st, onodes, oidx, _ = O.build_ploc(m.soa, 25)
    assert st == 0
    assert O.nodes_equal(nodes, onodes) is None
    np.testing.assert_array_equal(idx, oidx)
    ctx.cache_trim()


def _free_bytes(ctx):
    f, t = C.c_uint64(), C.c_uint64()
    assert ctx._lib.bvh_device_memory(ctx.h, C.byref(f), C.byref(t)) == 0
    return f.value, t.value


def test_freed_blocks_stay_counted_as_free(ctx):
    # Freed device blocks go back to the library's device pool, which keeps
    # the memory; bvh_device_memory counts it as free, before and after the
    # trim hands it back to the device.  Other processes on the device move
    # the figure too, so the tolerance is generous against a 168 MB mesh.
    ctx.cache_trim()
    before, total = _free_bytes(ctx)
    m = B.Mesh.uniform(406, 2_000_000)        # 72 MB of coordinates + 96 MB copy
    dm = ctx.upload(m)
    during, _ = _free_bytes(ctx)
    assert during <= before - 150_000_000
    dm.free()
    after_free, _ = _free_bytes(ctx)
    assert after_free >= before - 32_000_000
    ctx.cache_trim()
    after_trim, _ = _free_bytes(ctx)
    assert after_trim >= before - 32_000_000 and after_trim <= total
    # The pool's memory is reused: the same upload again fits where it was.
    dm = ctx.upload(m)
    assert _free_bytes(ctx)[0] <= before - 150_000_000
    dm.free()
