"""Meshes whose Morton codes are known in closed form (shared by the oracle
pins and the GPU parity tests).

Every triangle of these meshes is a point or has a box whose centroid is an
exact binary fraction, so the normalised centroid (lbvh_builder.cu:167-181:
`(c - min) / (ext > 1e-6 ? ext : 1)`) is exact and the expected cell is known
without evaluating the oracle.
"""
from __future__ import annotations

import numpy as np


def interleave30(x, y, z):
    """x in the most significant bit of each triple (morton3D_nt, lbvh_builder.cu:19-27)."""
    code = 0
    for b in range(10):
        code |= ((x >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((z >> b) & 1) << (3 * b)
    return code


def interleave60(x, y, z):
    """z in the most significant bit of each triple (morton3D_P, lbvh_plus_builder.cu:30-35)."""
    code = 0
    for b in range(20):
        code |= ((z >> b) & 1) << (3 * b + 2)
        code |= ((y >> b) & 1) << (3 * b + 1)
        code |= ((x >> b) & 1) << (3 * b)
    return code


def _points(pts):
    """Degenerate triangles: all three vertices at each point (box = the point)."""
    pts = np.asarray(pts, np.float32)
    soa = np.empty((9, len(pts)), np.float32)
    for v in range(3):
        soa[3 * v:3 * v + 3] = pts.T
    return soa


def cells_mesh(bits: int, seed: int = 0, count: int = 300, lo=(0.0, 0.0, 0.0), ext=(1.0, 1.0, 1.0)):
    """Point triangles at cell centres of a grid over the box [lo, lo + ext]
    (ext a power of two per axis, so every coordinate and every step of the
    normalisation is exact), two corner points that make the scene box exactly
    that box, and one asymmetric triangle whose box centroid differs from its
    vertex mean.  Returns (soa, expected codes)."""
    res = 1 << bits
    inter = interleave30 if bits == 10 else interleave60
    rng = np.random.default_rng(seed)
    ks = rng.integers(0, res, (count, 3))
    ks[:6] = [[0, 0, 0], [res - 1, res - 1, res - 1], [res - 1, 0, 0], [0, res - 1, 0], [0, 0, res - 1],
              [res // 2, res // 2 - 1, res // 2]]
    lo = np.asarray(lo, np.float64)
    ext = np.asarray(ext, np.float64)
    pts = [lo, lo + ext]
    want = [0, inter(res - 1, res - 1, res - 1)]
    for k in ks:
        pts.append(lo + ext * (k + 0.5) / res)
        want.append(inter(*[int(v) for v in k]))
    pts = np.asarray(pts, np.float64)
    assert np.all(pts == pts.astype(np.float32)), "fixture coordinates must be exact in float32"
    soa = _points(pts)
    # Asymmetric triangle: x vertices at 64, 512 and 100 units of 1/1024 of the
    # extent.  Box centroid 288 (the vertex mean, 225.3, would give cell 225);
    # y and z at fixed cell centres.
    u = ext / 1024.0
    tri = np.empty((9, 1), np.float64)
    for v, xu in enumerate((64, 512, 100)):
        tri[3 * v + 0, 0] = lo[0] + xu * u[0]
        tri[3 * v + 1, 0] = lo[1] + 10.5 * u[1]
        tri[3 * v + 2, 0] = lo[2] + 3.5 * u[2]
    assert np.all(tri == tri.astype(np.float32))
    soa = np.concatenate([soa, tri.astype(np.float32)], axis=1)
    scale = res // 1024
    want.append(inter(288 * scale, int(10.5 * scale), int(3.5 * scale)))
    return soa, np.array(want, dtype=np.uint64)


def flat_mesh(bits: int, z_top: float):
    """Two point triangles, (0,0,0) and (1,1,z_top): the z extent is z_top.
    With z_top <= 1e-6 the divisor is 1, so the second point's z cell is
    floor(z_top * 2^bits) (0 at 30 bits, 0 or 1 at 60); above it the divisor
    is the extent and that cell is the last one."""
    res = 1 << bits
    inter = interleave30 if bits == 10 else interleave60
    soa = _points([[0.0, 0.0, 0.0], [1.0, 1.0, z_top]])
    zt = np.float32(z_top)
    zcell = res - 1 if zt > np.float32(1e-6) else int(np.floor(np.float64(zt) * res))
    return soa, np.array([0, inter(res - 1, res - 1, zcell)], dtype=np.uint64)


FLAT_CASES = [2.0 ** -21, float(np.float32(1e-6)), 2.0 ** -19, 0.0]
