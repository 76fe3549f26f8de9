"""Size-independent properties of a BVH with one primitive per leaf (used by
the full-size GPU tests and checked on oracle trees on the CPU)."""
import numpy as np


def _sel_min(a, b):
    return np.where(b < a, b, a)


def _sel_max(a, b):
    return np.where(a < b, b, a)


def check_tree_properties(nodes, idx, tri_boxes):
    """Size-independent checks of a BVH over n primitives (one per leaf)."""
    n = idx.size
    assert nodes.size == 2 * n - 1
    left, right = nodes["left"], nodes["right"]
    internal = np.nonzero(left >= 0)[0]
    leaf = np.nonzero(left < 0)[0]
    assert internal.size == n - 1 and leaf.size == n
    # Children: in range, each non-root node referenced exactly once.
    ch = np.concatenate([left[internal], right[internal]]).astype(np.int64)
    assert ch.min() >= 0 and ch.max() < nodes.size
    refs = np.bincount(ch, minlength=nodes.size)
    assert refs[0] == 0 and np.all(refs[1:] == 1)
    # Reached from the root (no detached cycles): walk level by level.
    seen = 0
    frontier = np.array([0], np.int64)
    depth = 0
    while frontier.size:
        seen += frontier.size
        fi = frontier[left[frontier] >= 0]
        frontier = np.concatenate([left[fi], right[fi]]).astype(np.int64)
        depth += 1
        assert depth < 4096
    assert seen == nodes.size
    # Every primitive exactly once; leaves carry right = -1.
    prims = (left[leaf].astype(np.int64) & 0x7FFFFFFF)
    assert np.all(np.bincount(prims, minlength=n) == 1)
    assert np.all(right[leaf] == -1)
    np.testing.assert_array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))
    # Leaf boxes are the triangle boxes, bit for bit.
    np.testing.assert_array_equal(nodes["bmin"][leaf].view(np.uint32), tri_boxes[prims, 0:3].view(np.uint32))
    np.testing.assert_array_equal(nodes["bmax"][leaf].view(np.uint32), tri_boxes[prims, 4:7].view(np.uint32))
    # Every internal box = merge(left box, right box) with select min/max.
    lo = _sel_min(nodes["bmin"][left[internal]], nodes["bmin"][right[internal]])
    hi = _sel_max(nodes["bmax"][left[internal]], nodes["bmax"][right[internal]])
    np.testing.assert_array_equal(nodes["bmin"][internal].view(np.uint32), lo.view(np.uint32))
    np.testing.assert_array_equal(nodes["bmax"][internal].view(np.uint32), hi.view(np.uint32))
    return depth


def sah64(nodes, ct=3.0, ci=2.0):
    """SAH with one primitive per leaf: (Ct * sum of internal areas + Ci * sum of
    leaf areas) / root area, summed in float64 (order-independent to 1e-12)."""
    d = (nodes["bmax"] - nodes["bmin"]).astype(np.float64)
    a = 2.0 * (d[:, 0] * d[:, 1] + d[:, 1] * d[:, 2] + d[:, 2] * d[:, 0])
    internal = nodes["left"] >= 0
    return (ct * a[internal].sum() + ci * a[~internal].sum()) / a[0]
