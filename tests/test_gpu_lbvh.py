"""LBVH on the GPU (through the C ABI) against the CPU oracle: Morton codes,
sorted (code, index) keys, node array and index array must be bit-identical."""
import numpy as np
import pytest

import bvh_amd as B
import morton_fixtures as MF
import oracle as O
from test_oracle_pins import SIGNED_ZERO_BOXES, signed_zero_soa

pytestmark = pytest.mark.gpu


def _check_lbvh(ctx, mesh, radix_bits=0):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        codes = ctx.debug_morton(dm, B.ALGO_LBVH)
        ocodes = O.morton30(mesh.soa)
        np.testing.assert_array_equal(codes, ocodes.astype(np.uint64))
        times = ctx.build(dm, tree, algo=B.ALGO_LBVH, radix_bits=radix_bits, keep_keys=True)
        keys, _ = ctx.debug_sorted_keys(mesh.n)
        nodes, idx = tree.download()
    finally:
        tree.free()
        dm.free()
    onodes, oidx, okeys = O.build_lbvh(mesh.soa)
    np.testing.assert_array_equal(keys, okeys)
    np.testing.assert_array_equal(idx, oidx)
    err = O.nodes_equal(nodes, onodes)
    assert err is None, err
    return times


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 255, 256, 257, 2047, 2048, 2049, 4095, 4097, 65537])
def test_lbvh_sizes(ctx, n):
    _check_lbvh(ctx, B.Mesh.uniform(42 + n, n))


@pytest.mark.parametrize("bits", [8, 10])
def test_lbvh_radix_bits_invariant(ctx, bits):
    _check_lbvh(ctx, B.Mesh.blobs(5, 300_000), radix_bits=bits)


def test_lbvh_config0_1m(ctx):
    """1,000,000 uniform triangles, seed 42 (the CPU-runnable configuration)."""
    _check_lbvh(ctx, B.Mesh.uniform(42, 1_000_000))


def test_lbvh_blobs_4m(ctx):
    """4,000,000 blob triangles, seed 7 (the headline configuration)."""
    t = _check_lbvh(ctx, B.Mesh.blobs(7, 4_000_000))
    assert t["total_ms"] > 0


def test_lbvh_all_identical(ctx):
    n = 5000
    soa = np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32), (1, n))
    _check_lbvh(ctx, B.Mesh(soa))


def test_lbvh_flat_and_signed_zero(ctx):
    rng = np.random.default_rng(3)
    n = 20000
    soa = rng.random((9, n), dtype=np.float32)
    soa[2::3] = 0.0                      # flat in z: extent <= 1e-6 -> divisor 1.0
    soa[2, : n // 2] = -0.0              # mixed signed zeros
    soa[0, ::7] = -0.0
    soa[3, ::7] = 0.0
    _check_lbvh(ctx, B.Mesh(soa))


def test_lbvh_huge_offsets(ctx):
    m = B.Mesh.uniform(11, 50000)
    m.soa[:] = m.soa * np.float32(1000.0) + np.float32(1.0e6)
    _check_lbvh(ctx, m)


def test_lbvh_nonfinite_rejected(ctx):
    m = B.Mesh.uniform(1, 1000)
    m.soa[4, 17] = np.nan
    dm = ctx.upload(m)
    tree = ctx.tree()
    try:
        with pytest.raises(B.NativeError) as e:
            ctx.build(dm, tree, algo=B.ALGO_LBVH)
        assert e.value.status == 4
    finally:
        tree.free()
        dm.free()


def test_builder_interface_mirror(ctx):
    m = B.Mesh.uniform(42, 10000)
    b = B.LBVHBuilder(ctx)
    b.build(m)
    assert b.get_name() == "LBVH"
    onodes, oidx, _ = O.build_lbvh(m.soa)
    assert O.nodes_equal(b.get_nodes(), onodes) is None
    np.testing.assert_array_equal(b.get_indices(), oidx)
    assert b.get_last_build_time_ms() > 0
    text = b.get_timing_breakdown()
    names = [line.split(":")[0].strip() for line in text.splitlines()]
    assert names == ["Bounds/Centroids", "Morton Codes", "CUB Radix Sort", "Topology Build", "Bottom-Up Refit"]
    assert "Sort Pass 1" in b.stages()


@pytest.mark.parametrize("algo,bits", [(B.ALGO_LBVH, 10), (B.ALGO_PLOC, 10), (B.ALGO_LBVH_PLUS, 20)])
@pytest.mark.parametrize("lo,ext", [((0, 0, 0), (1, 1, 1)), ((2, -1, 0), (4, 2, 0.5)), ((-64, 8, -0.25), (32, 8, 2))])
def test_morton_known_cells_on_device(ctx, algo, bits, lo, ext):
    """Mesh-level Morton codes (scene box, box centroid, normalisation) equal
    closed-form cells, not only the oracle (lbvh_builder.cu:167-181)."""
    soa, want = MF.cells_mesh(bits, seed=bits + len(lo), lo=lo, ext=ext)
    dm = ctx.upload(B.Mesh(soa))
    try:
        np.testing.assert_array_equal(ctx.debug_morton(dm, algo), want)
    finally:
        dm.free()


@pytest.mark.parametrize("algo,bits", [(B.ALGO_LBVH, 10), (B.ALGO_LBVH_PLUS, 20)])
@pytest.mark.parametrize("z_top", MF.FLAT_CASES)
def test_morton_extent_rule_on_device(ctx, algo, bits, z_top):
    soa, want = MF.flat_mesh(bits, z_top)
    dm = ctx.upload(B.Mesh(soa))
    try:
        np.testing.assert_array_equal(ctx.debug_morton(dm, algo), want)
    finally:
        dm.free()


def test_signed_zero_leaf_boxes(ctx):
    """Leaf boxes keep the hand-derived signed-zero bits of the select rule."""
    soa = signed_zero_soa()
    n = soa.shape[1]
    nodes, idx = _build_download(ctx, B.Mesh(soa), B.ALGO_LBVH)
    leaves = nodes[n - 1:]
    for k in range(n):
        p = int(idx[k])
        _, lo_bits, hi_bits = SIGNED_ZERO_BOXES[p]
        assert int(leaves["bmin"][k, 0:1].view(np.uint32)[0]) == lo_bits, p
        assert int(leaves["bmax"][k, 0:1].view(np.uint32)[0]) == hi_bits, p


def _build_download(ctx, mesh, algo, **kw):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=algo, **kw)
        return tree.download()
    finally:
        tree.free()
        dm.free()
