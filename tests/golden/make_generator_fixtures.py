"""Writes tests/golden/generators.json: FNV-1a 64 checksums of the seeded
workload generators (include/bvh_scenes.h) at small sizes and at the sizes
the benchmark workloads use, plus the first triangles of each small mesh as
float32 bit patterns.  The file pins the generators: any change to a
generator changes every benchmark workload, and the tests that read this file
fail.  Run from the repository root after building:

    python tests/golden/make_generator_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import bvh_amd as B  # noqa: E402
from bvh_amd import _native as N  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generators.json")
FNV0 = 14695981039346656037


def fnv(a: np.ndarray, h: int = FNV0) -> int:
    a = np.ascontiguousarray(a)
    return int(N.host().bvh_fnv1a64(a.ctypes.data, a.nbytes, h))


def meshes():
    """(name, constructor) of every mesh the fixture covers."""
    out = [
        ("uniform_42_1000", lambda: B.Mesh.uniform(42, 1000)),
        ("uniform_42_1000000", lambda: B.Mesh.uniform(42, 1_000_000)),
        ("blobs_7_1000", lambda: B.Mesh.blobs(7, 1000)),
        ("blobs_7_4000000", lambda: B.Mesh.blobs(7, 4_000_000)),
        ("strands_99_100x160", lambda: B.Mesh.strands(99, 100, 160)),
        ("strands_99_2000x160", lambda: B.Mesh.strands(99, 2000, 160)),
    ]
    counts = B.batch_counts(11, 512, 1000, 200000)
    for i in range(4):
        out.append((f"batch_11_mesh{i}_{int(counts[i])}", lambda i=i: B.Mesh.batch_mesh(11, i, int(counts[i]))))
    return out


def rays():
    blob_box = B.Mesh.blobs(7, 4_000_000).bounds()
    return [
        ("camera_blobs_7_4000000_1024x768", lambda: B.rays_sphere_camera(blob_box, 1024, 768)),
        ("sphere_to_box_23_unit_1000", lambda: B.rays_sphere_to_box(23, np.array([0, 0, 0, 1, 1, 1], np.float32),
                                                                       1000)),
        ("uniform_cube_99_4096", lambda: B.rays_uniform_cube(99, 4096)),
    ]


def make() -> dict:
    d = {"fnv1a64_offset_basis": FNV0, "meshes": {}, "rays": {}, "first_triangles": {}}
    d["batch_counts_11_512_1000_200000"] = fnv(B.batch_counts(11, 512, 1000, 200000))
    for name, mk in meshes():
        m = mk()
        d["meshes"][name] = {"n": int(m.n), "checksum": m.checksum()}
        if m.n <= 400_000:
            d["first_triangles"][name] = [" ".join(f"{int(b):08x}" for b in m.soa[:, i].view(np.uint32)) for i in range(3)]
    for name, mk in rays():
        r = mk()
        d["rays"][name] = {"n": int(r.shape[0]), "checksum": fnv(r)}
    return d


if __name__ == "__main__":
    with open(OUT, "w") as f:
        json.dump(make(), f, indent=1, sort_keys=True)
        f.write("\n")
    print("wrote", OUT)
