"""Runs the reference's CLI (built unchanged against the drop-in headers by
tests/dropin_build.py) on the GPU: all builders plus a heatmap render."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_build", "dropin", "bvh_benchmark")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in CLI not built (needs the reference tree at build time)")
def test_reference_cli_runs_all_builders(tmp_path):
    out = tmp_path / "img"
    r = subprocess.run([EXE, "-n", "200000", "-a", "all", "--render", str(out), "--render-size", "256x192",
                        "--shading", "heatmap"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    text = r.stdout
    for name in ("LBVH", "LBVH+", "PLOC CUDA (r=10)", "PLOC CUDA (r=25)", "PLOC CUDA (r=100)"):
        assert name in text
    assert "FAILED" not in text
    assert text.count("SAH Cost") >= 5
    ppms = list(tmp_path.glob("img*.ppm"))
    assert len(ppms) == 5
    for p in ppms:
        assert p.read_bytes()[:2] == b"P6"


def test_obj_upload_and_reference_loader_agree(ctx, tmp_path):
    """bvh_mesh_upload_obj (parallel parse into pinned staging + DMA) builds the
    same tree as the parsed mesh uploaded from pageable memory; the reference
    CLI, whose loadOBJ is the reference's own loader.cpp, loads the same number
    of triangles and reports the same SAH for the LBVH tree."""
    import re
    import numpy as np
    import bvh_amd as B
    import oracle as O
    from test_obj_loader import _write_obj
    src = B.Mesh.blobs(61, 40_000)
    p = tmp_path / "blobs.obj"
    _write_obj(p, src)
    m = B.Mesh.from_obj(str(p))
    dm_a = ctx.upload_obj(str(p))
    dm_b = ctx.upload(m)
    ta, tb = ctx.tree(), ctx.tree()
    try:
        ctx.build(dm_a, ta, algo=B.ALGO_LBVH)
        ctx.build(dm_b, tb, algo=B.ALGO_LBVH)
        na, ia = ta.download()
        nb, ib = tb.download()
    finally:
        ta.free()
        tb.free()
        dm_a.free()
        dm_b.free()
    assert O.nodes_equal(na, nb) is None and np.array_equal(ia, ib)
    if not os.path.exists(EXE):
        pytest.skip("drop-in CLI not built")
    r = subprocess.run([EXE, "-i", str(p), "-a", "lbvh"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    loaded = re.search(r"Loaded (\d+) triangles", r.stdout)
    assert loaded and int(loaded.group(1)) == m.n
    sah = re.search(r"SAH Cost:\s+([0-9.]+)", r.stdout)
    assert sah and sah.group(1) == f"{O.sah(na):.2f}"


def test_reference_batch_harness_runs(tmp_path):
    """The reference's batch harness (batch_runner.cpp:395-514, batch_config.cpp
    JSON), compiled unchanged, runs every builder over an OBJ model and a
    random model on the GPU; the roofline report adds its columns."""
    import csv
    import json
    import bvh_amd as B
    from bvh_amd import batch_report
    from test_obj_loader import _write_obj
    if not os.path.exists(EXE):
        pytest.skip("drop-in CLI not built")
    obj = tmp_path / "m.obj"
    _write_obj(obj, B.Mesh.blobs(5, 30_000))
    out = tmp_path / "results.csv"
    cfg = {"iterations": 2, "warmup": True, "quiet": True, "output": str(out), "algorithms": ["all"],
           "ploc_radius": [10, 25],
           "models": [{"type": "obj", "path": str(obj), "name": "blobs30k"},
                      {"type": "random", "triangles": 100000}]}
    cj = tmp_path / "cfg.json"
    cj.write_text(json.dumps(cfg))
    r = subprocess.run([EXE, "--batch", str(cj)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    rows = list(csv.DictReader(open(out)))
    algos = {row["Algorithm"] for row in rows}
    assert {"LBVH", "LBVH+", "PLOC CUDA (r=10)", "PLOC CUDA (r=25)"} <= algos
    assert len(rows) == 2 * 2 * 4 and all(row["Status"] == "OK" for row in rows)
    assert {int(row["TriangleCount"]) for row in rows} == {30_000, 100_000}
    aug = batch_report.augment(rows)
    assert all(float(a["GBps"]) > 0 and 0 < float(a["FracHBM"]) < 1 for a in aug)
