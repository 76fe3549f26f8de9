"""The OpenMP oracle variants (CPU-baseline timing only) return exactly what
the serial oracle returns: trees, leaf order and PLOC iteration counts.  CPU
only."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O


@pytest.mark.parametrize("n", [2, 3, 1000, 60_000])
def test_lbvh_omp_equals_serial(n):
    m = B.Mesh.blobs(3, n)
    a, ia, _ = O.build_lbvh_omp(m)
    b, ib = O.build_lbvh(m)[:2]
    assert O.nodes_equal(a, b) is None and np.array_equal(ia, ib)


@pytest.mark.parametrize("n,rounds", [(2, 1), (7, 2), (1000, 1), (60_000, 3)])
def test_lbvh_plus_omp_equals_serial(n, rounds):
    m = B.Mesh.blobs(5, n)
    a, ia = O.build_lbvh_plus_omp(m, rounds)
    b, ib = O.build_lbvh_plus(m, rounds)[:2]
    assert O.nodes_equal(a, b) is None and np.array_equal(ia, ib)


def test_lbvh_plus_omp_strands():
    m = B.Mesh.strands(99, 300, 160)
    a, ia = O.build_lbvh_plus_omp(m, 2)
    b, ib = O.build_lbvh_plus(m, 2)[:2]
    assert O.nodes_equal(a, b) is None and np.array_equal(ia, ib)


@pytest.mark.parametrize("n,radius", [(2, 1), (5, 25), (1000, 4), (60_000, 25), (30_000, 100)])
def test_ploc_omp_equals_serial(n, radius):
    m = B.Mesh.uniform(9, n)
    st_a, a, ia, it_a = O.build_ploc_omp(m, radius)
    st_b, b, ib, iters = O.build_ploc(m, radius)
    assert st_a == st_b == 0
    assert O.nodes_equal(a, b) is None and np.array_equal(ia, ib)
    assert np.array_equal(it_a, iters)
