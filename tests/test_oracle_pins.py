"""Independent pins for the oracle functions that other tests only use as
inputs: SAH and depth statistics (evaluator.cpp:38-113), the per-triangle and
scene boxes (lbvh_builder.cu:47-66), the mesh-level Morton path with its
centroid and `ext > 1e-6` rules (lbvh_builder.cu:167-181,
lbvh_plus_builder.cu:276-290), the signed-zero box fixture, and the seeded
generators (RNG known answers and committed checksums).  Every expected value
here is a closed form, a hand computation or a published vector.  CPU only."""
import json
import os

import numpy as np
import pytest

import bvh_amd as B
import morton_fixtures as MF
import oracle as O
from bvh_amd import _native as N

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _node(lo, hi, left, right):
    n = np.zeros(1, dtype=O.NODE_DTYPE)
    n["bmin"] = lo
    n["bmax"] = hi
    n["left"] = left
    n["right"] = right
    return n


def _leaf(lo, hi, prim):
    return _node(lo, hi, np.int32(np.uint32(prim | 0x80000000).view(np.int32)), -1)


def _tree(*nodes):
    return np.concatenate(nodes)


# ---------------------------------------------------------------- SAH ------

def test_sah_identical_boxes_closed_form():
    # n identical triangles: every node has the root's box, so each internal
    # node contributes Ct and each leaf Ci: SAH = Ct*(n-1) + Ci*n, exactly.
    n = 1000
    tri = np.array([0, 0, 0, 1, 2, 0, 0, 2, 3], np.float32)   # box [0,1]x[0,2]x[0,3]
    soa = np.tile(tri[:, None], (1, n))
    nodes, _, _ = O.build_lbvh(soa)
    assert O.sah(nodes, 3.0, 2.0) == 3 * (n - 1) + 2 * n
    assert O.sah(nodes, 1.0, 1.0) == 2 * n - 1
    assert O.sah(nodes, 1.0, 0.0) == n - 1
    assert O.sah(nodes, 0.0, 1.0) == n


def test_sah_hand_computed_three_leaf_tree():
    # root [0,2]x[0,2]x[0,1]: SA 16; node1 [0,1]x[0,2]x[0,1]: SA 10;
    # leaves A [0,1]^3: 6, B [0,1]x[1,2]x{0}: 2, C [1,2]x[0,2]x{0}: 4.
    # SAH = Ct*(16+10)/16 + Ci*(6+2+4)/16  ->  6.375 at (3,2), 2.375 at (1,1).
    root = _node([0, 0, 0], [2, 2, 1], 1, 4)
    node1 = _node([0, 0, 0], [1, 2, 1], 2, 3)
    a = _leaf([0, 0, 0], [1, 1, 1], 0)
    b = _leaf([0, 1, 0], [1, 2, 0], 1)
    c = _leaf([1, 0, 0], [2, 2, 0], 2)
    nodes = _tree(root, node1, a, b, c)
    assert O.sah(nodes, 3.0, 2.0) == 6.375
    assert O.sah(nodes, 1.0, 1.0) == 2.375
    assert O.sah(nodes, 2.0, 3.0) == 2.0 * 26 / 16 + 3.0 * 12 / 16  # swapped costs differ
    # Same tree with a different node numbering: same cost.
    perm = _tree(_node([0, 0, 0], [2, 2, 1], 4, 1), c, b, a, _node([0, 0, 0], [1, 2, 1], 3, 2))
    assert O.sah(perm, 3.0, 2.0) == 6.375


def test_sah_two_leaves_and_degenerate_root():
    # Root SA 16, two leaves of SA 10 each: 3 + 2*(10/16)*2 = 5.5.
    nodes = _tree(_node([0, 0, 0], [2, 2, 1], 1, 2), _leaf([0, 0, 0], [1, 2, 1], 0), _leaf([1, 0, 0], [2, 2, 1], 1))
    assert O.sah(nodes, 3.0, 2.0) == 5.5
    # A root of zero area gives 0 (evaluator.cpp:44-45).
    flat = _tree(_node([0, 0, 0], [0, 0, 0], 1, 2), _leaf([0, 0, 0], [0, 0, 0], 0), _leaf([0, 0, 0], [0, 0, 0], 1))
    assert O.sah(flat, 3.0, 2.0) == 0.0


# ------------------------------------------------------- depth stats ------

def _complete_tree(k):
    # Heap layout: node i has children 2i+1, 2i+2; leaves at 2^k - 1 ...
    n_int = (1 << k) - 1
    parts = []
    for i in range(n_int):
        parts.append(_node([0, 0, 0], [1, 1, 1], 2 * i + 1, 2 * i + 2))
    for j in range(1 << k):
        parts.append(_leaf([0, 0, 0], [1, 1, 1], j))
    return _tree(*parts)


def _caterpillar(n):
    # Internal node i: left = leaf i, right = internal i+1; the last internal
    # node holds the last two leaves.  Leaf depths 1, 2, ..., n-1, n-1.
    parts = []
    for i in range(n - 1):
        right = i + 1 if i < n - 2 else (n - 1) + (n - 1)
        parts.append(_node([0, 0, 0], [1, 1, 1], (n - 1) + i, right))
    for j in range(n):
        parts.append(_leaf([0, 0, 0], [1, 1, 1], j))
    return _tree(*parts)


@pytest.mark.parametrize("k", [0, 1, 3, 10])
def test_depth_stats_complete_tree(k):
    nodes = _complete_tree(k)
    md, avg, leaves = O.depth_stats(nodes)
    assert (md, leaves) == (k, 1 << k)
    assert avg == k


@pytest.mark.parametrize("n", [2, 3, 50, 400])
def test_depth_stats_caterpillar(n):
    nodes = _caterpillar(n)
    md, avg, leaves = O.depth_stats(nodes)
    assert (md, leaves) == (n - 1, n)
    assert avg == pytest.approx(((n - 1) * n / 2 + (n - 1)) / n, rel=1e-12)


# ------------------------------------------------ boxes and Morton -------

def test_tri_boxes_and_scene_box_brute_force():
    rng = np.random.default_rng(5)
    soa = (rng.random((9, 2000), dtype=np.float32) * 4 - 2).astype(np.float32)
    v = soa.reshape(3, 3, -1)
    tb = O.tri_boxes(soa)
    np.testing.assert_array_equal(tb[:, 0:3], v.min(axis=0).T)
    np.testing.assert_array_equal(tb[:, 4:7], v.max(axis=0).T)
    sb = np.zeros(6, np.float32)
    O.lib().orc_scene_box(O._soa(soa), soa.shape[1], sb.ctypes.data)
    np.testing.assert_array_equal(sb[:3], v.min(axis=(0, 2)))
    np.testing.assert_array_equal(sb[3:], v.max(axis=(0, 2)))


# Signed zeros under select min/max: min(a,b) = (b < a) ? b : a and
# max(a,b) = (a < b) ? b : a keep the first operand when the values compare
# equal, and a box is min(v0, min(v1, v2)).  Hand-derived bit patterns of the
# x extent for each vertex order (v0x, v1x, v2x):
SIGNED_ZERO_BOXES = [
    ((-0.0, 0.0, 0.0), 0x80000000, 0x80000000),   # v0 wins both ties
    ((0.0, -0.0, -0.0), 0x00000000, 0x00000000),
    ((0.0, -0.0, 0.0), 0x00000000, 0x00000000),
    ((1.0, -0.0, 0.0), 0x80000000, 0x3F800000),   # inner min(v1, v2) keeps v1 = -0
    ((1.0, 0.0, -0.0), 0x00000000, 0x3F800000),   # ... and here v1 = +0
    ((-1.0, -0.0, 0.0), 0xBF800000, 0x80000000),  # inner max(v1, v2) keeps v1 = -0
    ((-1.0, 0.0, -0.0), 0xBF800000, 0x00000000),
]


def signed_zero_soa():
    n = len(SIGNED_ZERO_BOXES)
    soa = np.zeros((9, n), np.float32)
    for i, (xs, _, _) in enumerate(SIGNED_ZERO_BOXES):
        for v in range(3):
            soa[3 * v, i] = xs[v]
            soa[3 * v + 1, i] = 0.25 * i
            soa[3 * v + 2, i] = 1.0 + v
    return soa


def test_signed_zero_box_fixture():
    tb = O.tri_boxes(signed_zero_soa())
    for i, (_, lo_bits, hi_bits) in enumerate(SIGNED_ZERO_BOXES):
        assert int(tb[i, 0:1].view(np.uint32)[0]) == lo_bits, i
        assert int(tb[i, 4:5].view(np.uint32)[0]) == hi_bits, i


@pytest.mark.parametrize("bits", [10, 20])
@pytest.mark.parametrize("lo,ext", [((0, 0, 0), (1, 1, 1)), ((2, -1, 0), (4, 2, 0.5)), ((-64, 8, -0.25), (32, 8, 2))])
def test_mesh_morton_known_cells(bits, lo, ext):
    soa, want = MF.cells_mesh(bits, seed=bits + len(lo), lo=lo, ext=ext)
    got = O.morton30(soa).astype(np.uint64) if bits == 10 else O.morton60(soa)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("bits", [10, 20])
@pytest.mark.parametrize("z_top", MF.FLAT_CASES)
def test_mesh_morton_extent_rule(bits, z_top):
    soa, want = MF.flat_mesh(bits, z_top)
    got = O.morton30(soa).astype(np.uint64) if bits == 10 else O.morton60(soa)
    np.testing.assert_array_equal(got, want)


# ------------------------------------------------------- generators -------

M64 = (1 << 64) - 1


def _splitmix64(state):
    state = (state + 0x9E3779B97F4A7C15) & M64
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return state, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


def _xoshiro256ss(s):
    # Blackman & Vigna, "Scrambled linear pseudorandom number generators" (2021).
    result = (_rotl((s[1] * 5) & M64, 7) * 9) & M64
    t = (s[1] << 17) & M64
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = _rotl(s[3], 45)
    return result


def test_rng_known_answers():
    # Published SplitMix64 outputs for seed 0.
    st, outs = 0, []
    for _ in range(4):
        st, v = _splitmix64(st)
        outs.append(v)
    assert outs[:3] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    for seed in (0, 7, 11, 42, 99, 0xFFFFFFFFFFFFFFFF):
        st, s = seed, []
        for _ in range(4):
            st, v = _splitmix64(st)
            s.append(v)
        want = [_xoshiro256ss(s) for _ in range(64)]
        got = np.zeros(64, np.uint64)
        N.host().bvh_rng_u64(seed, 64, got.ctypes.data)
        assert [int(x) for x in got] == want, seed


def _fixtures():
    with open(os.path.join(GOLDEN, "generators.json")) as f:
        return json.load(f)


def test_generator_checksums_match_committed_table():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mkfix", os.path.join(GOLDEN, "make_generator_fixtures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    want = _fixtures()
    got = mk.make()
    assert got["batch_counts_11_512_1000_200000"] == want["batch_counts_11_512_1000_200000"]
    assert got["meshes"] == want["meshes"]
    assert got["rays"] == want["rays"]
    assert got["first_triangles"] == want["first_triangles"]


def test_generator_properties():
    m = B.Mesh.uniform(42, 20000)
    v = m.soa.reshape(3, 3, -1).astype(np.float64)
    spread = v.max(axis=0) - v.min(axis=0)
    assert spread.max() <= 0.01 + 1e-6
    assert v.min() >= -0.005 - 1e-6 and v.max() <= 1.005 + 1e-6

    n = 20000
    m = B.Mesh.blobs(7, n)
    v = m.soa.reshape(3, 3, -1).astype(np.float64)
    e1 = np.linalg.norm(v[1] - v[0], axis=0)
    e2 = np.linalg.norm(v[2] - v[0], axis=0)
    large = (e1 > 0.05)
    assert large.sum() == n - n * 95 // 100
    np.testing.assert_allclose(e1[large], 0.1, rtol=1e-5)
    np.testing.assert_allclose(e2[large], 0.1, rtol=1e-5)
    assert np.all(e1[~large] <= 0.006 * np.sqrt(3) + 1e-6)

    m = B.Mesh.strands(99, 50, 160)
    v = m.soa.reshape(3, 3, -1).astype(np.float64)
    seg = v[1] - v[0]
    w = v[2] - v[0]
    assert np.all(np.linalg.norm(seg, axis=0) <= 0.01 + 1e-6)
    np.testing.assert_allclose(np.linalg.norm(w, axis=0), 0.0005, rtol=1e-3)
    assert np.all(np.abs((seg * w).sum(axis=0)) <= 1e-8)
    assert v[:2].min() >= 0.0 and v[:2].max() <= 1.0
    # Consecutive segments of a strand share the joint bit for bit.
    s0 = m.soa[:, :160]
    np.testing.assert_array_equal(s0[0:3, 1:], s0[3:6, :-1])

    counts = B.batch_counts(11, 512, 1000, 200000)
    assert counts.min() >= 1000 and counts.max() <= 200000
