"""Tree files on the host (no device): the format of include/bvh_io.h round
trips, its header is checked, and a size query needs no buffers."""
import os
import struct

import numpy as np
import pytest

import bvh_amd as B


def _random_tree(n, seed):
    rng = np.random.default_rng(seed)
    nodes = np.zeros(2 * n - 1, dtype=B.NODE_DTYPE)
    nodes["bmin"] = rng.random((2 * n - 1, 3), dtype=np.float32)
    nodes["bmax"] = nodes["bmin"] + 1
    nodes["left"] = rng.integers(-2**31, 2**31 - 1, 2 * n - 1)
    nodes["right"] = rng.integers(-1, 2 * n, 2 * n - 1)
    return nodes, rng.permutation(n).astype(np.uint32)


def test_round_trip(tmp_path):
    nodes, idx = _random_tree(1000, 1)
    p = str(tmp_path / "t.bvh")
    B.tree_file_write(p, nodes, idx)
    n2, i2 = B.tree_file_read(p)
    assert n2.tobytes() == nodes.tobytes()
    np.testing.assert_array_equal(i2, idx)
    assert os.path.getsize(p) == 8 + 16 + 48 * nodes.size + 4 * idx.size


def test_header_layout(tmp_path):
    nodes, idx = _random_tree(3, 2)
    p = str(tmp_path / "t.bvh")
    B.tree_file_write(p, nodes, idx)
    raw = open(p, "rb").read()
    assert raw[:8] == b"BVHAMD01"
    assert struct.unpack("<4I", raw[8:24]) == (1, 5, 3, 0)
    assert raw[24:24 + 48] == nodes[:1].tobytes()


def test_rejects_foreign_file(tmp_path):
    p = str(tmp_path / "x.bvh")
    open(p, "wb").write(b"NOTATREE" + bytes(64))
    with pytest.raises(B.NativeError):
        B.tree_file_read(p)
    with pytest.raises(B.NativeError):
        B.tree_file_read(str(tmp_path / "missing.bvh"))


def test_rejects_truncated_and_oversized(tmp_path):
    nodes, idx = _random_tree(100, 3)
    p = str(tmp_path / "t.bvh")
    B.tree_file_write(p, nodes, idx)
    raw = open(p, "rb").read()
    short = str(tmp_path / "short.bvh")
    open(short, "wb").write(raw[:-4])
    with pytest.raises(B.NativeError):
        B.tree_file_read(short)
    # A header claiming 2^31 nodes over a tiny file is refused before any allocation.
    big = str(tmp_path / "big.bvh")
    open(big, "wb").write(raw[:8] + struct.pack("<4I", 1, 1 << 31, 1 << 30, 0) + raw[24:24 + 96])
    with pytest.raises(B.NativeError):
        B.tree_file_read(big)


def test_validate_tree_structure():
    import oracle as O
    from bvh_amd import _native as N
    m = B.Mesh.uniform(9, 500)
    for nodes, idx in (O.build_lbvh(m.soa)[:2], O.build_ploc(m.soa, 8)[1:3]):
        v = N.host().bvh_tree_validate
        assert v(nodes.ctypes.data, nodes.size, idx.ctypes.data, idx.size) == 0
        bad = nodes.copy()
        bad["right"][0] = nodes.size            # child out of range
        assert v(bad.ctypes.data, bad.size, idx.ctypes.data, idx.size) == 8
        bad = nodes.copy()
        leaf = int(np.nonzero(nodes["left"] < 0)[0][0])
        bad["left"][leaf] = np.int32(np.uint32(0x80000000 | 500).view(np.int32))   # primitive out of range
        assert v(bad.ctypes.data, bad.size, idx.ctypes.data, idx.size) == 8
        assert v(nodes.ctypes.data, nodes.size - 2, idx.ctypes.data, idx.size) == 8   # wrong node count
        badi = idx.copy()
        badi[3] = 500
        assert v(nodes.ctypes.data, nodes.size, badi.ctypes.data, idx.size) == 8


def _obj_text(boxes):
    """The wireframe text for boxes (min, max), built by hand: 8 corners
    (bottom face counter-clockwise, then top), two closed face loops and four
    vertical edges per box, indices from 1, floats in the stream's default
    (%g, 6 significant digits) format."""
    out = []
    o = 1
    for lo, hi in boxes:
        c = [(lo[0], lo[1], lo[2]), (hi[0], lo[1], lo[2]), (hi[0], hi[1], lo[2]), (lo[0], hi[1], lo[2]),
             (lo[0], lo[1], hi[2]), (hi[0], lo[1], hi[2]), (hi[0], hi[1], hi[2]), (lo[0], hi[1], hi[2])]
        out += ["v %s %s %s" % tuple("%g" % x for x in p) for p in c]
        out.append(f"l {o} {o + 1} {o + 2} {o + 3} {o}")
        out.append(f"l {o + 4} {o + 5} {o + 6} {o + 7} {o + 4}")
        out += [f"l {o + e} {o + 4 + e}" for e in range(4)]
        o += 8
    return "\n".join(out) + "\n"


def test_obj_export_matches_hand_written_text(tmp_path):
    nodes = np.zeros(3, dtype=B.NODE_DTYPE)
    nodes["bmin"] = [[0, 0, 0], [0.25, -1.5, 2], [1.125, 0, 3e-7]]
    nodes["bmax"] = [[2, 1, 4], [1, 0.5, 3], [2, 1, 4]]
    nodes["left"] = [1, np.int32(np.uint32(0x80000000).view(np.int32)), np.int32(np.uint32(0x80000001).view(np.int32))]
    nodes["right"] = [2, -1, -1]
    from bvh_amd import _native as N
    p = str(tmp_path / "all.obj")
    assert N.host().bvh_nodes_export_obj(nodes.ctypes.data, 3, p.encode(), 0) == 0
    boxes = [(nodes["bmin"][i], nodes["bmax"][i]) for i in range(3)]
    assert open(p).read() == _obj_text(boxes)
    p = str(tmp_path / "leaves.obj")
    assert N.host().bvh_nodes_export_obj(nodes.ctypes.data, 3, p.encode(), 1) == 0
    assert open(p).read() == _obj_text(boxes[1:])


def test_obj_export_leaves_of_ploc_tree(tmp_path):
    """PLOC keeps leaves first (root swapped to 0): the leaf bit selects them,
    where the reference's index test would pick internal nodes."""
    import oracle as O
    from bvh_amd import _native as N
    m = B.Mesh.uniform(4, 200)
    _, nodes, _, _ = O.build_ploc(m.soa, 8)
    p = str(tmp_path / "ploc.obj")
    assert N.host().bvh_nodes_export_obj(nodes.ctypes.data, nodes.size, p.encode(), 1) == 0
    text = open(p).read()
    leaf = np.nonzero(nodes["left"] < 0)[0]
    assert text.count("\nl ") + text.startswith("l ") == 6 * leaf.size
    assert text == _obj_text([(nodes["bmin"][i], nodes["bmax"][i]) for i in leaf])
    wrong = [i for i in range(nodes.size) if i >= (nodes.size + 1) // 2 - 1]   # the reference's test
    assert set(wrong) != set(leaf.tolist())
