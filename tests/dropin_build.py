"""Builds the reference's own consumer programs (main.cpp, batch_runner.cpp,
batch_config.cpp, loader.cpp, evaluator.cpp — compiled unchanged, read where
they lie in the reference tree) against this repository's drop-in headers and
links them with libbvhhost.so + libbvhhip.so.  The result, _build/dropin/
bvh_benchmark, is the reference's CLI running on the MI355X engine.

The reference sources include their headers by relative path ("../include/x.h",
"cuda/lbvh_builder.cuh"); a throw-away tree of symlinks in _build/dropin/tree
routes those includes to this repository's headers (common.h, bvh_node.h,
mesh.h, bvh_builder.h, render_engine.h, bvh_amd/cuda/*.cuh) and to the
reference's own consumer headers (evaluator.h, batch_config.h, batch_runner.h).
"""
from __future__ import annotations

import os
import shutil
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("BVH_REFERENCE", "/root/reference")
OUT = os.path.join(ROOT, "_build", "dropin")
SOURCES = ["main.cpp", "batch_runner.cpp", "batch_config.cpp", "loader.cpp", "evaluator.cpp"]
OUR_HEADERS = ["common.h", "bvh_node.h", "mesh.h", "bvh_builder.h", "render_engine.h"]
REF_HEADERS = ["evaluator.h", "batch_config.h", "batch_runner.h"]
OUR_BUILDERS = ["lbvh_builder.cuh", "lbvh_plus_builder.cuh", "ploc_builder.cuh"]


def available() -> bool:
    return all(os.path.exists(os.path.join(REF, "src", s)) for s in SOURCES)


def build() -> str:
    tree = os.path.join(OUT, "tree")
    shutil.rmtree(tree, ignore_errors=True)
    os.makedirs(os.path.join(tree, "src", "cuda"))
    os.makedirs(os.path.join(tree, "include"))
    os.makedirs(os.path.join(tree, "src", "host"))
    for s in SOURCES:
        os.symlink(os.path.join(REF, "src", s), os.path.join(tree, "src", s))
    for h in REF_HEADERS:
        os.symlink(os.path.join(REF, "include", h), os.path.join(tree, "include", h))
    for h in OUR_HEADERS:
        os.symlink(os.path.join(ROOT, "include", h), os.path.join(tree, "include", h))
    # Drop-in builder headers include "../../include/bvh_builder.h" and
    # "../host/hip_builder.hpp" relative to src/cuda.
    for h in OUR_BUILDERS:
        os.symlink(os.path.join(ROOT, "bvh_amd", "cuda", h), os.path.join(tree, "src", "cuda", h))
    os.symlink(os.path.join(ROOT, "bvh_amd", "host", "hip_builder.hpp"),
               os.path.join(tree, "src", "host", "hip_builder.hpp"))
    lib = os.path.join(ROOT, "bvh_amd")
    exe = os.path.join(OUT, "bvh_benchmark")
    cmd = ["g++", "-std=c++17", "-O2", "-I", os.path.join(tree, "include")]
    cmd += [os.path.join(tree, "src", s) for s in SOURCES]
    cmd += ["-o", exe, f"-L{lib}", "-lbvhhost", "-lbvhhip", f"-Wl,-rpath,{lib}", "-Wl,-rpath,$ORIGIN/../../bvh_amd",
            "-L/opt/rocm/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    shutil.rmtree(tree, ignore_errors=True)
    return exe


if __name__ == "__main__":
    print(build())
