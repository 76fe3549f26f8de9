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
