"""Pins the CPU oracle's LBVH path (Morton codes, sort order, Karras topology,
refit) against known answers, closed forms and brute force.  CPU only."""
import numpy as np
import pytest

import oracle as O


def interleave30(x, y, z):
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def interleave60(x, y, z):
    code = 0
    for b in range(20):
        code |= ((z >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((x >> b) & 1) << (3 * b)
    return code


def test_expand_known_answers():
    L = O.lib()
    assert L.orc_expand10(1023) == 0x09249249
    assert L.orc_expand10(1) == 1
    assert L.orc_expand10(2) == 8
    for v in range(1024):
        want = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert L.orc_expand10(v) == want
    rng = np.random.default_rng(1)
    for v in rng.integers(0, 1 << 20, 2000):
        want = sum(((int(v) >> b) & 1) << (3 * b) for b in range(20))
        assert L.orc_expand21(int(v)) == want


def test_morton_extremes_and_cells():
    L = O.lib()
    assert L.orc_morton30_point(1.0, 1.0, 1.0) == 0x3FFFFFFF
    assert L.orc_morton30_point(0.0, 0.0, 0.0) == 0
    assert L.orc_morton60_point(1.0, 1.0, 1.0) == 0x0FFFFFFFFFFFFFFF
    # Clamping below zero and above one.
    assert L.orc_morton30_point(-5.0, 2.0, 0.0) == interleave30(0, 1023, 0)
    # Exact cell boundaries k/1024 quantise to k (closed form).
    for k in [0, 1, 2, 511, 512, 1000, 1023]:
        p = np.float32(k) / np.float32(1024)
        assert L.orc_morton30_point(p, 0.0, 0.0) == interleave30(k, 0, 0)
        assert L.orc_morton30_point(0.0, p, 0.0) == interleave30(0, k, 0)
        assert L.orc_morton30_point(0.0, 0.0, p) == interleave30(0, 0, k)
    for k in [0, 1, 2, 524287, 524288, 1048575]:
        p = np.float32(k) / np.float32(1048576)
        assert L.orc_morton60_point(p, 0.0, 0.0) == interleave60(k, 0, 0)
        assert L.orc_morton60_point(0.0, 0.0, p) == interleave60(0, 0, k)


def test_morton_random_points_vs_bit_loop():
    L = O.lib()
    rng = np.random.default_rng(2)
    pts = rng.random((3000, 3), dtype=np.float32)
    for x, y, z in pts:
        q = [int(min(max(np.float32(c) * np.float32(1024.0), 0.0), 1023.0)) for c in (x, y, z)]
        assert L.orc_morton30_point(x, y, z) == interleave30(*q)
        q = [int(min(max(np.float32(c) * np.float32(1048576.0), 0.0), 1048575.0)) for c in (x, y, z)]
        assert L.orc_morton60_point(x, y, z) == interleave60(*q)


def test_karras_paper_figure3():
    # Karras 2012, Figure 3: keys 00001 00010 00100 00101 10011 11000 11001 11110.
    codes = np.array([0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110], np.uint32)
    idx = np.arange(8, dtype=np.uint32)
    left, right, parent = O.karras30(codes, idx)
    leaf = lambda k: 7 + k  # noqa: E731
    want = {0: (3, 4), 1: (leaf(0), leaf(1)), 2: (leaf(2), leaf(3)), 3: (1, 2), 4: (leaf(4), 5),
            5: (6, leaf(7)), 6: (leaf(5), leaf(6))}
    for i, (l, r) in want.items():
        assert (left[i], right[i]) == (l, r), i


def _topdown(keys):
    """Textbook top-down radix-tree construction over sorted unique integer keys;
    the node numbering follows Karras 2012 Sec. 4 (root 0, left child = split,
    right child = split + 1, leaves offset by n-1)."""
    n = len(keys)
    left = np.zeros(max(n - 1, 0), np.int64)
    right = np.zeros(max(n - 1, 0), np.int64)
    stack = [(0, n - 1, 0)]
    while stack:
        first, last, node = stack.pop()
        d = (keys[first] ^ keys[last]).bit_length()  # highest differing bit + 1
        split = first
        for j in range(first, last):
            if ((keys[first] ^ keys[j]).bit_length()) < d:
                split = j
        lc = split if split != first else (n - 1 + split)
        rc = split + 1 if split + 1 != last else (n - 1 + split + 1)
        left[node], right[node] = lc, rc
        if split != first:
            stack.append((first, split, split))
        if split + 1 != last:
            stack.append((split + 1, last, split + 1))
    return left, right


@pytest.mark.parametrize("n,dup", [(2, 0), (3, 0), (17, 0), (300, 0), (300, 1), (1500, 1)])
def test_karras_matches_topdown_recursion(n, dup):
    rng = np.random.default_rng(n + dup)
    hi = 16 if dup else (1 << 30)
    codes = np.sort(rng.integers(0, hi, n).astype(np.uint32))
    idx = rng.permutation(n).astype(np.uint32)
    # Stable (code, index) order.
    order = np.lexsort((idx, codes))
    codes, idx = codes[order], idx[order]
    keys = [(int(c) << 32) | int(i) for c, i in zip(codes, idx)]
    tl, tr = _topdown(keys)
    left, right, _ = O.karras30(codes, idx)
    np.testing.assert_array_equal(left.astype(np.int64), tl)
    np.testing.assert_array_equal(right.astype(np.int64), tr)


def _uniform_soa(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((3, n), dtype=np.float32)
    soa = np.empty((9, n), np.float32)
    for v in range(3):
        soa[3 * v:3 * v + 3] = c + (rng.random((3, n), dtype=np.float32) * np.float32(0.01) - np.float32(0.005))
    return soa


def _check_tree(nodes, idx, soa):
    n = soa.shape[1]
    assert nodes.size == 2 * n - 1
    # Every primitive appears in exactly one leaf; sorted order == indices.
    leaves = nodes[n - 1:]
    assert np.all(leaves["left"] < 0)
    np.testing.assert_array_equal((leaves["left"] & 0x7FFFFFFF).astype(np.uint32), idx)
    assert sorted(idx.tolist()) == list(range(n))
    # Brute-force refit: each internal box = union over its subtree's triangles.
    tb = O.tri_boxes(soa)
    lo = np.full((2 * n - 1, 3), np.inf, np.float32)
    hi = np.full((2 * n - 1, 3), -np.inf, np.float32)
    for k in range(n):
        lo[n - 1 + k] = tb[idx[k], :3]
        hi[n - 1 + k] = tb[idx[k], 4:7]
    order = []
    st = [0]
    while st:
        i = st.pop()
        order.append(i)
        if nodes["left"][i] >= 0:
            st += [nodes["left"][i], nodes["right"][i]]
    for i in reversed(order):
        if nodes["left"][i] >= 0:
            l, r = nodes["left"][i], nodes["right"][i]
            lo[i] = np.minimum(lo[l], lo[r])
            hi[i] = np.maximum(hi[l], hi[r])
    assert len(order) == 2 * n - 1
    np.testing.assert_array_equal(nodes["bmin"], lo)
    np.testing.assert_array_equal(nodes["bmax"], hi)


@pytest.mark.parametrize("n", [1, 2, 3, 64, 65, 1000])
def test_lbvh_oracle_invariants(n):
    soa = _uniform_soa(n, 7 + n)
    nodes, idx, keys = O.build_lbvh(soa)
    _check_tree(nodes, idx, soa)
    codes = O.morton30(soa)
    order = np.lexsort((np.arange(n), codes))
    np.testing.assert_array_equal(idx, order.astype(np.uint32))
    np.testing.assert_array_equal(keys, (codes[order].astype(np.uint64) << 32) | order.astype(np.uint64))


def test_lbvh_oracle_identical_triangles():
    n = 200
    soa = np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32), (1, n))
    nodes, idx, _ = O.build_lbvh(soa)
    _check_tree(nodes, idx, soa)
    np.testing.assert_array_equal(idx, np.arange(n, dtype=np.uint32))


def test_omp_build_matches_serial():
    soa = _uniform_soa(20000, 3)
    a, ia, _ = O.build_lbvh(soa)
    b, ib, _ = O.build_lbvh_omp(soa)
    assert O.nodes_equal(a, b) is None
    np.testing.assert_array_equal(ia, ib)


def test_sah_cost_decomposition():
    # With one primitive per leaf SAH(Ct, Ci) = Ct*I + Ci*L (I, L: internal and
    # leaf area sums over the root area).  A property, not a pin: the closed
    # forms in test_oracle_pins.py pin the SAH itself.
    soa = _uniform_soa(5000, 9)
    nodes, _, _ = O.build_lbvh(soa)
    i_sum = O.sah(nodes, 1.0, 0.0)
    l_sum = O.sah(nodes, 0.0, 1.0)
    assert abs(O.sah(nodes, 1.0, 1.0) - (i_sum + l_sum)) <= 1e-4 * (i_sum + l_sum)
    assert abs(O.sah(nodes, 3.0, 2.0) - (3 * i_sum + 2 * l_sum)) <= 1e-4 * (3 * i_sum + 2 * l_sum)
