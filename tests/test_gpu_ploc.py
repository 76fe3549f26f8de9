"""PLOC on the GPU against the oracle: node array (topology, deterministic
numbering, root swap, boxes) and index array bit-identical."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _check_ploc(ctx, mesh, radius):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        t = ctx.build(dm, tree, algo=B.ALGO_PLOC, radius=radius)
        nodes, idx = tree.download()
    finally:
        tree.free()
        dm.free()
    st, onodes, oidx, iters = O.build_ploc(mesh.soa, radius)
    assert st == 0
    np.testing.assert_array_equal(idx, oidx)
    err = O.nodes_equal(nodes, onodes)
    assert err is None, err
    assert t["ploc_iterations"] == len(iters)
    return t


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 255, 256, 257, 1024, 2047, 2048, 2049, 2050, 4095, 4097, 65537])
def test_ploc_sizes(ctx, n):
    _check_ploc(ctx, B.Mesh.uniform(100 + n, n), 25)


@pytest.mark.parametrize("radius", [1, 2, 10, 16, 100])
def test_ploc_radii(ctx, radius):
    _check_ploc(ctx, B.Mesh.blobs(3 + radius, 120_000), radius)


def test_ploc_blobs_1m_r25(ctx):
    _check_ploc(ctx, B.Mesh.blobs(7, 1_000_000), 25)


@pytest.mark.parametrize("radius", [100, 200])
def test_ploc_blobs_1m_wide_radius(ctx, radius):
    """Large-m neighbour search (three clusters per lane): symmetric pair
    evaluation up to r = 128, the one-sided scan above it."""
    _check_ploc(ctx, B.Mesh.blobs(11, 1_000_000), radius)


@pytest.mark.parametrize("n,radius", [(20000, 25), (300_000, 25), (300_000, 100)])
def test_ploc_signed_zero(ctx, n, radius):
    """Mixed signed zeros and flat boxes: the neighbour search evaluates each
    pair once from the lower cluster's side (300K clusters reach that path),
    which is exact only if SA(a u b) and SA(b u a) have the same bits."""
    rng = np.random.default_rng(n + radius)
    soa = rng.random((9, n), dtype=np.float32)
    soa[2::3] = 0.0
    soa[2, : n // 2] = -0.0
    soa[0, ::7] = -0.0
    soa[3, ::7] = 0.0
    _check_ploc(ctx, B.Mesh(soa), radius)


def test_ploc_identical_triangles(ctx):
    n = 2600  # one merge per iteration: exercises ~550 global iterations and the LDS tail
    soa = np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32), (1, n))
    _check_ploc(ctx, B.Mesh(soa), 2)


def test_ploc_builder_mirror(ctx):
    m = B.Mesh.uniform(9, 30000)
    b = B.PLOCBuilder(25, ctx)
    b.build(m)
    assert b.get_name() == "PLOC CUDA (r=25)"
    st, onodes, oidx, _ = O.build_ploc(m.soa, 25)
    assert O.nodes_equal(b.get_nodes(), onodes) is None
    np.testing.assert_array_equal(b.get_indices(), oidx)


def test_ploc_bad_radius(ctx):
    m = B.Mesh.uniform(1, 100)
    dm = ctx.upload(m)
    tree = ctx.tree()
    try:
        with pytest.raises(B.NativeError) as e:
            ctx.build(dm, tree, algo=B.ALGO_PLOC, radius=0)
        assert e.value.status == 1
    finally:
        tree.free()
        dm.free()
