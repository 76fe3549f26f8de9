"""Builds the native libraries of bvh_amd in-tree.

  libbvhhip.so   gfx950 kernels + the C ABI of include/bvh_hip.h (hipcc)
  libbvhhost.so  host C++ above the C ABI: BVHBuilder classes, renderer,
                 scene generators (g++; links libbvhhip.so)

Usage: python -m bvh_amd.build   (or bvh_amd.build.build_all())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("BVH_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every float op is one IEEE op in source order (parity with
# the CPU oracle).  Packed fp32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", f"-I{os.path.join(ROOT, 'include')}",
]
HOST_FLAGS = ["-O2", "-fPIC", "-pthread", "-std=c++17", "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
              f"-I{HERE}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _hip_obj(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", out])
    return out


def _host_obj(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([CXX, *HOST_FLAGS, "-c", src, "-o", out])
    return out


def build_all(verbose: bool = False) -> None:
    os.makedirs(OBJ, exist_ok=True)
    hip_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    host_srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) \
        if os.path.isdir(HOST) else []
    jobs = min(8, max(1, os.cpu_count() or 1))
    with cf.ThreadPoolExecutor(jobs) as ex:
        hip_objs = list(ex.map(_hip_obj, hip_srcs))
        host_objs = list(ex.map(_host_obj, host_srcs))
    lib = os.path.join(HERE, "libbvhhip.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *hip_objs, "-o", lib + ".tmp"])
    os.replace(lib + ".tmp", lib)
    if host_objs:
        hlib = os.path.join(HERE, "libbvhhost.so")
        _run([CXX, "-shared", "-fPIC", *host_objs, "-o", hlib + ".tmp", f"-L{HERE}", "-lbvhhip",
              "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined", "-L/opt/rocm/lib", "-lamdhip64", "-pthread"])
        os.replace(hlib + ".tmp", hlib)
    if verbose:
        print("built", lib)


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


if __name__ == "__main__":
    build_all(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle()
