"""The reference's own consumer programs (main.cpp, batch_runner.cpp,
batch_config.cpp, loader.cpp, evaluator.cpp) compile unchanged against the
drop-in headers and link against the MI355X libraries.  CPU only: builds when
the reference tree is present, then runs `--help` (no GPU work)."""
import os
import subprocess

import pytest

import dropin_build


@pytest.mark.skipif(not dropin_build.available(), reason="reference tree absent")
def test_reference_cli_builds_against_dropin_headers():
    exe = dropin_build.build()
    assert os.path.exists(exe)
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "lbvh, lbvh+, ploc, all" in r.stdout
