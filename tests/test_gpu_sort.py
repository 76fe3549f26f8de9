"""The onesweep radix sort on the GPU against numpy's stable argsort."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 5, 4095, 4096, 4097, 100_000, 1_000_003])
@pytest.mark.parametrize("bits", [8, 10])
def test_sort_pairs_random(ctx, n, bits):
    rng = np.random.default_rng(n * 31 + bits)
    keys = rng.integers(0, 1 << 62, n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    k, v, _ = ctx.sort_pairs(keys, vals, 0, 64, bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("lo,hi", [(32, 62), (0, 60), (8, 20)])
def test_sort_partial_bits_is_stable(ctx, lo, hi):
    rng = np.random.default_rng(lo + hi)
    n = 300_000
    keys = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    k, _, _ = ctx.sort_pairs(keys, None, lo, hi, 0)
    mask = np.uint64(((1 << (hi - lo)) - 1) << lo)
    order = np.argsort(keys & mask, kind="stable")
    np.testing.assert_array_equal(k, keys[order])


def test_sort_few_distinct_digits(ctx):
    rng = np.random.default_rng(9)
    n = 500_000
    keys = (rng.integers(0, 3, n).astype(np.uint64) << np.uint64(40)) | np.arange(n, dtype=np.uint64)
    k, _, _ = ctx.sort_pairs(keys, None, 32, 62, 10)
    order = np.argsort(keys >> np.uint64(32), kind="stable")
    np.testing.assert_array_equal(k, keys[order])


@pytest.mark.parametrize("n", [1, 8191, 8192, 8193, 1_000_003, 2_499_999, 2_500_000, 2_506_753, 3_000_001])
@pytest.mark.parametrize("bits", [8, 10])
def test_sort_keys_only_tile_edges(ctx, n, bits):
    """Key-only sorts at tile edges: 8192-key tiles, and from 2.5M keys with
    8-bit digits the wide 16384-key tiles (2,506,753 = 153 wide tiles + 1)."""
    rng = np.random.default_rng(n * 7 + bits)
    keys = rng.integers(0, 1 << 62, n, dtype=np.uint64)
    k, _, _ = ctx.sort_pairs(keys, None, 0, 62, bits)
    np.testing.assert_array_equal(k, np.sort(keys, kind="stable"))
