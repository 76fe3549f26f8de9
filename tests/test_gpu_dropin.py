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
