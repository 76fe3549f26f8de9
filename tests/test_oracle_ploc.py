"""Pins the oracle's PLOC against a naive Python transcription of the reference
loop (ploc_builder.cu:173-264,489-548 + root swap :563-582) on small meshes,
and checks tree invariants and the progress property.  CPU only."""
import numpy as np
import pytest

import oracle as O

F = np.float32


def fmin(a, b):
    return b if b < a else a


def fmax(a, b):
    return b if a < b else a


def merge(a, b):
    return tuple(fmin(a[k], b[k]) for k in range(3)) + tuple(fmax(a[3 + k], b[3 + k]) for k in range(3))


def area(b):
    dx, dy, dz = F(b[3] - b[0]), F(b[4] - b[1]), F(b[5] - b[2])
    s = F(F(dx * dy) + F(dy * dz))
    s = F(s + F(dz * dx))
    return F(F(2.0) * s)


def naive_ploc(soa, radius):
    n = soa.shape[1]
    tb = O.tri_boxes(soa)
    codes = O.morton30(soa)
    order = np.lexsort((np.arange(n), codes))
    boxes = [tuple(F(x) for x in (tb[p, 0], tb[p, 1], tb[p, 2], tb[p, 4], tb[p, 5], tb[p, 6])) for p in order]
    nodes = [(boxes[i], int(order[i]) | -0x80000000, -1) for i in range(n)]
    ids = list(range(n))
    while len(boxes) > 1:
        m = len(boxes)
        nn = []
        for i in range(m):
            best, bj = F(np.finfo(np.float32).max), -1
            for j in range(max(0, i - radius), min(m - 1, i + radius) + 1):
                if j == i:
                    continue
                d = area(merge(boxes[i], boxes[j]))
                if d < best:
                    best, bj = d, j
                elif d == best and bj == -1:
                    bj = j
            nn.append(bj)
        nb, ni = [], []
        for i in range(m):
            j = nn[i]
            mutual = j != -1 and nn[j] == i
            if mutual and i < j:
                u = merge(boxes[i], boxes[j])
                nodes.append((u, ids[i], ids[j]))
                nb.append(u)
                ni.append(len(nodes) - 1)
            elif mutual and j < i:
                pass
            else:
                nb.append(boxes[i])
                ni.append(ids[i])
        assert len(nb) < m, "no progress"
        boxes, ids = nb, ni
    if len(nodes) > 1:
        root = len(nodes) - 1
        nodes[0], nodes[root] = nodes[root], nodes[0]
        fixed = []
        for b, l, r in nodes:
            if l >= 0:
                l = root if l == 0 else (0 if l == root else l)
                r = root if r == 0 else (0 if r == root else r)
            fixed.append((b, l, r))
        nodes = fixed
    return nodes, order


@pytest.mark.parametrize("n,radius", [(1, 1), (2, 1), (3, 1), (17, 2), (64, 3), (150, 5), (300, 25), (257, 100)])
def test_ploc_oracle_matches_naive_transcription(n, radius):
    rng = np.random.default_rng(n * 7 + radius)
    c = rng.random((3, n), dtype=np.float32)
    soa = np.empty((9, n), np.float32)
    for v in range(3):
        soa[3 * v:3 * v + 3] = c + rng.random((3, n), dtype=np.float32) * F(0.05)
    st, nodes, idx, iters = O.build_ploc(soa, radius)
    assert st == 0
    want, order = naive_ploc(soa, radius)
    np.testing.assert_array_equal(idx, order.astype(np.uint32))
    assert nodes.size == len(want) == 2 * n - 1
    for k, (b, l, r) in enumerate(want):
        assert nodes["left"][k] == l and nodes["right"][k] == r, k
        np.testing.assert_array_equal(nodes["bmin"][k], np.array(b[:3], np.float32))
        np.testing.assert_array_equal(nodes["bmax"][k], np.array(b[3:], np.float32))


def test_ploc_oracle_invariants_and_progress():
    rng = np.random.default_rng(5)
    n = 5000
    soa = rng.random((9, n), dtype=np.float32)
    st, nodes, idx, iters = O.build_ploc(soa, 25)
    assert st == 0
    # Cluster count strictly decreases every iteration.
    assert np.all(np.diff(iters) < 0) and iters[0] == n
    # Full binary tree over every primitive, root at 0, each node referenced once.
    refs = np.concatenate([nodes["left"][nodes["left"] >= 0], nodes["right"][nodes["left"] >= 0]])
    assert sorted(refs.tolist()) == list(range(1, 2 * n - 1))
    leaves = nodes[nodes["left"] < 0]
    assert sorted((leaves["left"] & 0x7FFFFFFF).tolist()) == list(range(n))
    # Every internal box is the union of its children.
    internal = np.nonzero(nodes["left"] >= 0)[0]
    l, r = nodes["left"][internal], nodes["right"][internal]
    np.testing.assert_array_equal(nodes["bmin"][internal], np.minimum(nodes["bmin"][l], nodes["bmin"][r]))
    np.testing.assert_array_equal(nodes["bmax"][internal], np.maximum(nodes["bmax"][l], nodes["bmax"][r]))


def test_ploc_oracle_identical_boxes_terminate():
    n = 300
    soa = np.tile(np.array([[0.1], [0.2], [0.3], [0.4], [0.2], [0.3], [0.1], [0.5], [0.3]], np.float32), (1, n))
    st, nodes, idx, iters = O.build_ploc(soa, 4)
    assert st == 0 and len(iters) == n - 1
    want, _ = naive_ploc(soa, 4)
    assert [int(x) for x in nodes["left"]] == [w[1] for w in want]
