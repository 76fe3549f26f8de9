"""bvh_build_lbvh_noalloc (the no-allocation LBVH build) through the C ABI:
two calls in a row, the second on a smaller mesh, each bit-identical to
bvh_build's tree for the same mesh, with no change in free device memory
across the calls; the size limit and bad arguments are rejected."""
import ctypes as C

import numpy as np
import pytest

import bvh_amd as B
from bvh_amd import _native as N
import oracle as O

pytestmark = pytest.mark.gpu


class _Dev:
    """Device copies of a mesh's nine SoA arrays plus output buffers."""

    def __init__(self, ctx, mesh):
        self.ctx, self.lib = ctx, N.hip()
        n = mesh.n
        self.n = n
        self.ptrs = []
        self.soa = (C.c_void_p * 9)()
        for k in range(9):
            p = self._alloc(n * 4)
            ctx._check(self.lib.bvh_copy_h2d(ctx.h, p, mesh.soa[k].ctypes.data, n * 4), "h2d")
            self.soa[k] = p
        self.nodes = self._alloc((2 * n - 1) * 48)
        self.indices = self._alloc(n * 4)

    def _alloc(self, nbytes):
        p = C.c_void_p()
        self.ctx._check(self.lib.bvh_device_alloc(self.ctx.h, nbytes, C.byref(p)), "alloc")
        self.ptrs.append(p)
        return p

    def build(self):
        return self.lib.bvh_build_lbvh_noalloc(self.ctx.h, self.soa, self.n, self.nodes, self.indices)

    def download(self):
        nodes = np.zeros(2 * self.n - 1, dtype=B.NODE_DTYPE)
        idx = np.zeros(self.n, dtype=np.uint32)
        self.ctx._check(self.lib.bvh_copy_d2h(self.ctx.h, nodes.ctypes.data, self.nodes, nodes.nbytes), "d2h")
        self.ctx._check(self.lib.bvh_copy_d2h(self.ctx.h, idx.ctypes.data, self.indices, idx.nbytes), "d2h")
        return nodes, idx

    def free(self):
        for p in self.ptrs:
            self.lib.bvh_device_free(self.ctx.h, p)
        self.ptrs = []


def _reference(ctx, mesh):
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=B.ALGO_LBVH)
        return tree.download()
    finally:
        tree.free()
        dm.free()


def _free_bytes(ctx):
    free, total = C.c_uint64(), C.c_uint64()
    ctx._check(N.hip().bvh_device_memory(ctx.h, C.byref(free), C.byref(total)), "memory info")
    return free.value


def test_noalloc_matches_bvh_build_twice(ctx):
    meshes = [B.Mesh.blobs(31, N.NOALLOC_MAX_TRIANGLES), B.Mesh.uniform(32, 37_123)]
    refs = [_reference(ctx, m) for m in meshes]
    devs = [_Dev(ctx, m) for m in meshes]
    try:
        before = _free_bytes(ctx)
        for d in devs:
            assert d.build() == 0, N.hip().bvh_last_error(ctx.h).decode()
        after = _free_bytes(ctx)
        for d, (rnodes, ridx), m in zip(devs, refs, meshes):
            nodes, idx = d.download()
            np.testing.assert_array_equal(idx, ridx)
            assert O.nodes_equal(nodes, rnodes) is None
        assert after == before
        # Again in the other order: the scratch keeps the last call's contents.
        for d in reversed(devs):
            assert d.build() == 0
        nodes, idx = devs[0].download()
        assert O.nodes_equal(nodes, refs[0][0]) is None
        np.testing.assert_array_equal(idx, refs[0][1])
    finally:
        for d in devs:
            d.free()


def test_noalloc_matches_oracle_small(ctx):
    for n in (1, 2, 3, 1025, 4097):
        m = B.Mesh.uniform(700 + n, n)
        d = _Dev(ctx, m)
        try:
            assert d.build() == 0
            nodes, idx = d.download()
        finally:
            d.free()
        onodes, oidx, _ = O.build_lbvh(m.soa)
        np.testing.assert_array_equal(idx, oidx)
        assert O.nodes_equal(nodes, onodes) is None


def test_noalloc_rejects_bad_arguments(ctx):
    lib = N.hip()
    m = B.Mesh.uniform(5, 100)
    d = _Dev(ctx, m)
    try:
        assert lib.bvh_build_lbvh_noalloc(ctx.h, d.soa, N.NOALLOC_MAX_TRIANGLES + 1, d.nodes, d.indices) == 1
        assert b"limit" in lib.bvh_last_error(ctx.h)
        assert lib.bvh_build_lbvh_noalloc(ctx.h, d.soa, 0, d.nodes, d.indices) == 1
        assert lib.bvh_build_lbvh_noalloc(ctx.h, None, 100, d.nodes, d.indices) == 1
        m.soa[3, 50] = np.inf
        ctx._check(lib.bvh_copy_h2d(ctx.h, d.soa[3], m.soa[3].ctypes.data, 400), "h2d")
        assert lib.bvh_build_lbvh_noalloc(ctx.h, d.soa, 100, d.nodes, d.indices) == 4
    finally:
        d.free()
