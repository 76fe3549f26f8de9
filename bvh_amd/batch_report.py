"""Adds algorithmic-bandwidth columns to the reference batch harness's CSV.

The reference harness (src/batch_runner.cpp:219-251, run unchanged over these
builders) writes one row per model, algorithm and iteration with BuildTimeMs
and TriangleCount.  This appends, per row:

  AlgBytesPerTri  algorithmic HBM bytes of one build per triangle:
                  LBVH 344 (bounds 36, Morton 44, four 8-bit sort passes of
                  16, topology 24, leaves 92, refit 84);
                  LBVH+ (one treelet round, as the reference builds it) 344
                  with the 60-bit key + payload sort (eight passes of 24
                  instead of four of 16) + 912 per internal node for the round;
                  PLOC r: init 224 (bounds, Morton, sort, leaf gather and
                  write) + 124 per live cluster per iteration + 48 per new
                  node, with the cluster visits taken as n (the first
                  iteration only), so the PLOC figure is a lower bound;
  GBps            AlgBytesPerTri x TriangleCount / BuildTimeMs;
  FracHBM         GBps / 8000 (MI355X HBM3E peak).

  python -m bvh_amd.batch_report batch_results.csv [out.csv]
"""
from __future__ import annotations

import csv
import sys

HBM_PEAK_GBS = 8000.0
LBVH_BYTES = 36 + 44 + 4 * 16 + 24 + 92 + 84


def bytes_per_tri(algorithm: str) -> float:
    a = algorithm.strip()
    if a == "LBVH":
        return float(LBVH_BYTES)
    if a == "LBVH+":
        return float(LBVH_BYTES - 4 * 16 + 8 * 24 + 912)
    if a.startswith("PLOC"):
        return float(224 + 124 + 48)
    raise ValueError(f"unknown algorithm {algorithm!r}")


def augment(rows):
    out = []
    for r in rows:
        r = dict(r)
        try:
            n = float(r["TriangleCount"])
            ms = float(r["BuildTimeMs"])
            b = bytes_per_tri(r["Algorithm"])
        except (KeyError, ValueError):
            r["AlgBytesPerTri"] = r["GBps"] = r["FracHBM"] = "N/A"
            out.append(r)
            continue
        gbs = b * n / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        r["AlgBytesPerTri"] = f"{b:.0f}"
        r["GBps"] = f"{gbs:.1f}"
        r["FracHBM"] = f"{gbs / HBM_PEAK_GBS:.4f}"
        out.append(r)
    return out


def main(argv):
    src = argv[1]
    dst = argv[2] if len(argv) > 2 else src.replace(".csv", "_roofline.csv")
    with open(src, newline="") as f:
        rd = csv.DictReader(f)
        fields = list(rd.fieldnames or [])
        rows = augment(list(rd))
    with open(dst, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields + ["AlgBytesPerTri", "GBps", "FracHBM"])
        w.writeheader()
        w.writerows(rows)
    print(dst)


if __name__ == "__main__":
    main(sys.argv)
