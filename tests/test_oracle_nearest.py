"""Pins the oracle's nearest-first traversal (the engine's BVH_TRACE_NEAREST
order): its hit is the least (t, triangle index) pair over every triangle,
checked against an all-triangle brute force; its visit counts against
hand-built trees whose answers are derived by hand below.  CPU only."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

LEAF = -0x80000000  # int32 image of 0x80000000


def mesh_of(tris):
    """Mesh from a list of triangles ((x,y,z) x 3)."""
    a = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    soa = np.empty((9, a.shape[0]), np.float32)
    for v in range(3):
        for k in range(3):
            soa[3 * v + k] = a[:, v, k]
    return B.Mesh(soa)


def tree_of(shape, mesh):
    """Node array for a nested tuple of primitive ids (a leaf is an int, an
    internal node a pair), root at 0, boxes the unions of triangle boxes."""
    soa = mesh.soa
    nodes = []

    def box_of(p):
        v = soa[:, p].reshape(3, 3)
        return v.min(axis=0), v.max(axis=0)

    def add(s):
        i = len(nodes)
        nodes.append(None)
        if isinstance(s, int):
            mn, mx = box_of(s)
            nodes[i] = (mn, mx, LEAF | s, -1)
            return i, mn, mx
        li, lmn, lmx = add(s[0])
        ri, rmn, rmx = add(s[1])
        mn, mx = np.minimum(lmn, rmn), np.maximum(lmx, rmx)
        nodes[i] = (mn, mx, li, ri)
        return i, mn, mx

    add(shape)
    out = np.zeros(len(nodes), dtype=B.NODE_DTYPE)
    for i, (mn, mx, l, r) in enumerate(nodes):
        out[i]["bmin"], out[i]["bmax"], out[i]["left"], out[i]["right"] = mn, mx, l, r
    return out


def tri_at_z(z, x0=0.0, size=1.0):
    """Triangle in the plane z (normal +z) covering [x0, x0+size] x [0, size]."""
    return [(x0, 0.0, z), (x0 + size, 0.0, z), (x0, size, z)]


RAY_DOWN_Z = np.array([[0.25, 0.25, 10.0, 0.0, 0.0, -1.0]], np.float32)  # origin above, looking down -z


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (50, 3), (400, 4), (3000, 5)])
def test_nearest_matches_brute_force(n, seed):
    m = B.Mesh.uniform(seed, n)
    rays = B.rays_sphere_to_box(seed, m.bounds(), 4000)
    bt, bp = O.trace_brute(m.soa, rays)   # strict '<' in ascending index order: the least (t, index)
    for build in (lambda s: O.build_lbvh(s)[0], lambda s: O.build_ploc(s, 8)[1],
                  lambda s: O.build_lbvh_plus(s, 2)[0]):
        nodes = build(m.soa)
        t, prim, cnt = O.trace(nodes, m.soa, rays, nearest=True)
        np.testing.assert_array_equal(prim, bp)
        np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
        assert np.all(cnt[:, 0] == cnt[:, 1])
        assert np.all(cnt[:, 0] % 2 == 1)  # root + two per expanded node


def test_nearest_matches_brute_force_on_blobs_with_camera_rays():
    m = B.Mesh.blobs(7, 20000)
    rays = B.rays_sphere_camera(m.bounds(), 96, 72)
    bt, bp = O.trace_brute(m.soa, rays)
    for nodes in (O.build_lbvh(m.soa)[0], O.build_ploc(m.soa, 25)[1]):
        t, prim, cnt = O.trace(nodes, m.soa, rays, nearest=True)
        np.testing.assert_array_equal(prim, bp)
        np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
        rt, rp, rc = O.trace(nodes, m.soa, rays)
        # Same hits as the reference order here; it never visits more on average.
        np.testing.assert_array_equal(prim, rp)
        assert cnt[:, 0].mean() <= rc[:, 0].mean() and cnt[:, 2].mean() <= rc[:, 2].mean()


def test_counts_single_leaf_and_root_miss():
    m = mesh_of([tri_at_z(0.0)])
    nodes = tree_of(0, m)
    t, p, c = O.trace(nodes, m.soa, RAY_DOWN_Z, nearest=True)
    assert p[0] == 0 and t[0] == np.float32(10.0)
    assert c[0].tolist() == [1, 1, 1]             # the root (a leaf): one box, one triangle
    miss = np.array([[5.0, 5.0, 10.0, 0.0, 0.0, -1.0]], np.float32)
    t, p, c = O.trace(nodes, m.soa, miss, nearest=True)
    assert p[0] == -1 and t[0] == np.float32(1e30)
    assert c[0].tolist() == [1, 1, 0]             # root box missed: nothing entered


def test_counts_near_child_first_culls_the_far_one():
    # Triangle 0 at z=0 (far), triangle 1 at z=5 (near), the far one as the
    # left child.  Reference order: root, both children counted (3 boxes),
    # the far leaf first (hit at t=10), then the near one (entry 5 < 10: hit
    # at t=5): 2 triangle tests.  Nearest order: the near leaf first (hit at
    # t=5), the far leaf's entry 10 > 5: skipped, 1 triangle test.
    m = mesh_of([tri_at_z(0.0), tri_at_z(5.0)])
    nodes = tree_of((0, 1), m)
    t, p, c = O.trace(nodes, m.soa, RAY_DOWN_Z, nearest=True)
    assert p[0] == 1 and t[0] == np.float32(5.0) and c[0].tolist() == [3, 3, 1]
    t, p, c = O.trace(nodes, m.soa, RAY_DOWN_Z)
    assert p[0] == 1 and t[0] == np.float32(5.0) and c[0].tolist() == [3, 3, 2]


def test_counts_four_leaves_in_depth_order():
    # Four stacked triangles at z = 0, 2, 4, 6 in a balanced tree
    # ((z0, z2), (z4, z6)).  Nearest order: root (1) + its two children (2);
    # the right pair is nearer (entry 4 < 0's entry 10): expanded (+2); z6
    # first (hit t=4), z4's entry 6 > 4 skipped; back at the left pair, its
    # entry 8 > 4: skipped.  Boxes 5, triangles 1.  Reference order: left
    # pair first (+2): z0 (t=10), z2 (entry 8 < 10, t=8); right pair (entry 4
    # < 8, +2): z4 (entry 6 < 8, t=6), z6 (entry 4 < 6, t=4).  Boxes 7,
    # triangles 4.
    m = mesh_of([tri_at_z(0.0), tri_at_z(2.0), tri_at_z(4.0), tri_at_z(6.0)])
    nodes = tree_of(((0, 1), (2, 3)), m)
    t, p, c = O.trace(nodes, m.soa, RAY_DOWN_Z, nearest=True)
    assert p[0] == 3 and t[0] == np.float32(4.0) and c[0].tolist() == [5, 5, 1]
    t, p, c = O.trace(nodes, m.soa, RAY_DOWN_Z)
    assert p[0] == 3 and t[0] == np.float32(4.0) and c[0].tolist() == [7, 7, 4]


def test_equal_t_takes_the_smaller_index_whatever_the_tree_order():
    # Triangles 2 and 5 are the same triangle; the others are off the ray.
    # The reference order returns whichever leaf it meets first (strict '<');
    # the nearest order returns index 2 in both trees, as the brute force does.
    far = [tri_at_z(0.0, x0=50.0 + 3 * k) for k in range(6)]
    tris = list(far)
    tris[2] = tri_at_z(1.0)
    tris[5] = tri_at_z(1.0)
    m = mesh_of(tris)
    bt, bp = O.trace_brute(m.soa, RAY_DOWN_Z)
    assert bp[0] == 2
    for shape, ref_first in (((((5, 0), (1, 3)), (2, 4)), 5), ((((2, 0), (1, 3)), (5, 4)), 2)):
        nodes = tree_of(shape, m)
        t, p, _ = O.trace(nodes, m.soa, RAY_DOWN_Z, nearest=True)
        assert p[0] == 2 and t[0] == np.float32(9.0)
        t, p, _ = O.trace(nodes, m.soa, RAY_DOWN_Z)
        assert p[0] == ref_first
