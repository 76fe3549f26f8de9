"""Clusters entering each PLOC iteration for a generated workload, from the
OpenMP oracle (whose trees the GPU tests match bit for bit, so the GPU runs
the same iterations), with the algorithmic bytes they imply: 124 B per
cluster per iteration (search 32 + 4, merge and compaction 44 read + 36
write, scan 8) plus 48 B per new node.  Test infrastructure.

  python oracle/ploc_iterations.py blobs 7 4000000 25
  python oracle/ploc_iterations.py strands 99 200000x160 100
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import bvh_amd as B  # noqa: E402
import oracle as O  # noqa: E402


def main():
    kind, seed, size, radius = sys.argv[1], int(sys.argv[2]), sys.argv[3], int(sys.argv[4])
    if kind == "blobs":
        mesh = B.Mesh.blobs(seed, int(size))
        desc = f"blobs seed {seed}, {int(size):,} triangles"
    elif kind == "strands":
        s, k = (int(x) for x in size.split("x"))
        mesh = B.Mesh.strands(seed, s, k)
        desc = f"strands seed {seed}, {s:,} x {k} segments"
    else:
        raise SystemExit("mesh kind: blobs or strands")
    st, _, _, m = O.build_ploc_omp(mesh, radius, max_iters=100_000)
    if st != 0:
        raise SystemExit(f"PLOC status {st}")
    total = int(m.astype(int).sum())
    n = mesh.n
    print(json.dumps({"workload": f"PLOC r={radius} on {desc}", "triangles": n, "radius": radius,
                      "iterations": int(len(m)), "clusters_per_iteration": [int(x) for x in m],
                      "sum_clusters": total, "bytes_per_cluster_iteration": 124, "bytes_per_new_node": 48,
                      "algorithmic_bytes": 124 * total + 48 * (n - 1)}))


if __name__ == "__main__":
    main()
