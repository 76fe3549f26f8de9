"""Pins the oracle's traversal (render_engine.cu:194-356) against an
all-triangle brute-force closest hit, and the seeded ray generators.  CPU only."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (50, 3), (400, 4), (3000, 5)])
def test_trace_matches_brute_force(n, seed):
    m = B.Mesh.uniform(seed, n)
    rays = B.rays_sphere_to_box(seed, m.bounds(), 4000)
    for build in (lambda s: O.build_lbvh(s)[0], lambda s: O.build_ploc(s, 8)[1],
                  lambda s: O.build_lbvh_plus(s, 2)[0]):
        nodes = build(m.soa)
        t, prim, cnt = O.trace(nodes, m.soa, rays)
        bt, bp = O.trace_brute(m.soa, rays)
        np.testing.assert_array_equal(prim, bp)
        np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
        assert np.all(cnt[:, 0] == cnt[:, 1])          # every visited node gets one AABB test
        assert np.all(cnt[:, 0] % 2 == 1)              # root + two per expanded node
        t2, p2, c2 = O.trace(nodes, m.soa, rays, omp=True)
        np.testing.assert_array_equal(p2, prim)
        np.testing.assert_array_equal(c2, cnt)


def test_camera_rays_are_unit_and_aimed():
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    rays = B.rays_sphere_camera(box, 64, 48)
    assert rays.shape == (64 * 48, 6)
    d = rays[:, 3:].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # The image centre looks at the box centre.
    eye = rays[0, :3].astype(np.float64)
    fwd = np.array([0.5, 0.5, 0.5]) - eye
    fwd /= np.linalg.norm(fwd)
    centre = rays[24 * 64 + 32, 3:].astype(np.float64)
    assert np.dot(centre, fwd) > 0.999


def test_generators_are_seeded_and_stable():
    a = B.Mesh.blobs(7, 10000)
    b = B.Mesh.blobs(7, 10000)
    c = B.Mesh.blobs(8, 10000)
    assert a.checksum() == b.checksum() != c.checksum()
    counts = B.batch_counts(11, 512, 1000, 200000)
    assert counts.min() >= 1000 and counts.max() <= 200000
    assert 25_000 < counts.mean() < 50_000      # log-uniform mean (199000 / ln 200 = 37.6K)
