"""Pins the oracle's LBVH+ treelet optimisation (lbvh_plus_builder.cu:71-247):
the rebuilt treelet attains the minimum total internal-node area over every
binary tree on its leaves (explicit enumeration, not the DP), leaves and
parent links are preserved, and SAH never increases.  CPU only."""
import itertools

import numpy as np
import pytest

import oracle as O


def _soa(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.random((3, n), dtype=np.float32) * np.float32(spread)
    soa = np.empty((9, n), np.float32)
    for v in range(3):
        soa[3 * v:3 * v + 3] = c + rng.random((3, n), dtype=np.float32) * np.float32(0.05)
    return soa


def _area(lo, hi):
    d = hi.astype(np.float64) - lo.astype(np.float64)
    return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0])


def _all_trees(items):
    """Every full binary tree (as nested tuples) over the given leaf labels."""
    items = tuple(items)
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    # Split into (left containing `first`, right) to enumerate each unordered tree once.
    for r in range(0, len(rest)):
        for left_rest in itertools.combinations(rest, r):
            right = tuple(x for x in rest if x not in left_rest)
            for lt in _all_trees((first,) + left_rest):
                for rt in _all_trees(right):
                    yield (lt, rt)


def _tree_cost(t, boxes):
    """(sum of internal-node areas, lo, hi) of an explicit tree."""
    if not isinstance(t, tuple):
        lo, hi = boxes[t]
        return 0.0, lo, hi
    cl, lol, hil = _tree_cost(t[0], boxes)
    cr, lor, hir = _tree_cost(t[1], boxes)
    lo, hi = np.minimum(lol, lor), np.maximum(hil, hir)
    return cl + cr + _area(lo, hi), lo, hi


def test_tree_enumeration_count():
    assert sum(1 for _ in _all_trees(range(7))) == 10395   # (2*7-3)!!
    assert sum(1 for _ in _all_trees(range(4))) == 15


def _subtree_internal_area(nodes, root):
    tot, st = 0.0, [root]
    while st:
        i = st.pop()
        if nodes["left"][i] >= 0:
            tot += _area(nodes["bmin"][i], nodes["bmax"][i])
            st += [nodes["left"][i], nodes["right"][i]]
    return tot


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (5, 3), (7, 4), (7, 5), (7, 6), (7, 7)])
def test_treelet_dp_is_optimal_over_all_trees(n, seed):
    soa = _soa(n, seed)
    nodes, idx, _ = O.build_lbvh(soa)
    # Parent links of the plain LBVH.
    parents = np.full(nodes.size, 0xFFFFFFFF, np.uint32)
    for i in range(nodes.size):
        if nodes["left"][i] >= 0:
            parents[nodes["left"][i]] = i
            parents[nodes["right"][i]] = i
    opt, par2 = O.optimize_treelet(nodes, parents, 0)
    leaf_ids = list(range(n - 1, 2 * n - 1))
    boxes = {i: (nodes["bmin"][i], nodes["bmax"][i]) for i in leaf_ids}
    best = min(_tree_cost(t, boxes)[0] for t in _all_trees(leaf_ids))
    got = _subtree_internal_area(opt, 0)
    assert abs(got - best) <= 1e-6 * best
    # Leaves untouched, every leaf still reachable once, parent links consistent.
    np.testing.assert_array_equal(opt[n - 1:], nodes[n - 1:])
    reach, st = [], [0]
    while st:
        i = st.pop()
        reach.append(i)
        if opt["left"][i] >= 0:
            for c in (opt["left"][i], opt["right"][i]):
                assert par2[c] == i
                st.append(c)
    assert sorted(reach) == list(range(2 * n - 1))


@pytest.mark.parametrize("n", [1, 2, 3, 8, 64, 500, 3000])
def test_lbvh_plus_sah_never_worse(n):
    soa = _soa(n, n, spread=3.0)
    base, bidx, _ = O.build_lbvh(soa)
    prev = O.sah(base, 1.0, 1.0)
    for rounds in (1, 2, 3):
        nodes, idx, codes = O.build_lbvh_plus(soa, rounds)
        assert nodes.size == 2 * n - 1
        np.testing.assert_array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))
        leaves = nodes[n - 1:]
        np.testing.assert_array_equal((leaves["left"] & 0x7FFFFFFF).astype(np.uint32), idx)
        s = O.sah(nodes, 1.0, 1.0)
        assert s <= prev * (1 + 1e-5)
        prev = s
    # 60-bit z-major codes sorted ascending with index ties.
    order = np.lexsort((np.arange(n), O.morton60(soa)))
    np.testing.assert_array_equal(idx, order.astype(np.uint32))


def test_tree_property_checker_on_oracle_trees():
    """The property checker used at full size accepts every builder's tree and
    rejects a tree with one swapped child box or a duplicated primitive."""
    import bvh_amd as B
    from tree_props import check_tree_properties, sah64
    m = B.Mesh.strands(99, 40, 160)
    tb = O.tri_boxes(m.soa)
    trees = [O.build_lbvh(m.soa)[:2], O.build_lbvh_plus(m.soa, 3)[:2], O.build_ploc(m.soa, 100)[1:3]]
    for nodes, idx in trees:
        check_tree_properties(nodes, idx, tb)
    s0 = sah64(O.build_lbvh_plus(m.soa, 0)[0])
    s3 = sah64(trees[1][0])
    assert s3 < s0
    nodes, idx = trees[0]
    bad = nodes.copy()
    bad["bmin"][3] = bad["bmin"][4]
    with pytest.raises(AssertionError):
        check_tree_properties(bad, idx, tb)
    bad = nodes.copy()
    n = idx.size
    bad["left"][n] = bad["left"][n + 1]
    with pytest.raises(AssertionError):
        check_tree_properties(bad, idx, tb)
