"""LBVH+ on the GPU against the oracle: 60-bit sorted keys, the optimised node
array and the index array bit-identical; SAH (evaluator semantics) equal."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _check(ctx, mesh, rounds):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        codes = ctx.debug_morton(dm, B.ALGO_LBVH_PLUS)
        np.testing.assert_array_equal(codes, O.morton60(mesh.soa))
        t = ctx.build(dm, tree, algo=B.ALGO_LBVH_PLUS, rounds=rounds, keep_keys=True)
        keys, vals = ctx.debug_sorted_keys(mesh.n)
        nodes, idx = tree.download()
    finally:
        tree.free()
        dm.free()
    onodes, oidx, ocodes = O.build_lbvh_plus(mesh.soa, rounds)
    np.testing.assert_array_equal(keys, ocodes)
    np.testing.assert_array_equal(vals, oidx)
    np.testing.assert_array_equal(idx, oidx)
    err = O.nodes_equal(nodes, onodes)
    assert err is None, err
    s_gpu, s_ref = O.sah(nodes), O.sah(onodes)
    assert abs(s_gpu - s_ref) <= 1e-5 * abs(s_ref)
    return t


@pytest.mark.parametrize("n", [1, 2, 3, 7, 8, 63, 64, 65, 257, 2047, 2048, 2049, 4097, 65537])
def test_lbvh_plus_sizes(ctx, n):
    _check(ctx, B.Mesh.uniform(500 + n, n), 1)


@pytest.mark.parametrize("rounds", [0, 1, 2, 3])
def test_lbvh_plus_rounds(ctx, rounds):
    _check(ctx, B.Mesh.blobs(21, 200_000), rounds)


def _signed_zero_mesh(seed, n):
    rng = np.random.default_rng(seed)
    soa = rng.random((9, n), dtype=np.float32)
    soa[2::3] = 0.0                      # flat in z
    soa[2, : n // 2] = -0.0              # mixed signed zeros
    soa[0, ::7] = -0.0
    soa[3, ::7] = 0.0
    return B.Mesh(soa)


def test_lbvh_plus_signed_zero(ctx):
    """Mixed signed zeros: treelet subset unions are select folds in leaf
    order, so the zero signs of the optimised boxes match the oracle's bits."""
    _check(ctx, _signed_zero_mesh(5, 20000), 2)


def test_lbvh_plus_blobs_1m_three_rounds(ctx):
    _check(ctx, B.Mesh.blobs(7, 1_000_000), 3)


def test_lbvh_plus_identical(ctx):
    n = 3000
    soa = np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32), (1, n))
    _check(ctx, B.Mesh(soa), 2)
