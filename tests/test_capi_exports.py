"""The C-ABI libraries load without a GPU and export every function that
include/*.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bvh_\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


@pytest.mark.parametrize("header,lib", [("bvh_hip.h", "libbvhhip.so"), ("bvh_scenes.h", "libbvhhost.so"),
                                        ("bvh_server.h", "libbvhhost.so"), ("bvh_io.h", "libbvhhost.so")])
def test_exports(header, lib):
    names = declared(header)
    assert len(names) >= 5
    path = os.path.join(ROOT, "bvh_amd", lib)
    assert os.path.exists(path), f"{path} not built"
    L = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL) if lib == "libbvhhip.so" else None
    if L is None:
        ctypes.CDLL(os.path.join(ROOT, "bvh_amd", "libbvhhip.so"), mode=ctypes.RTLD_GLOBAL)
        L = ctypes.CDLL(path)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_status_strings_without_device():
    L = ctypes.CDLL(os.path.join(ROOT, "bvh_amd", "libbvhhip.so"))
    L.bvh_status_string.restype = ctypes.c_char_p
    assert L.bvh_status_string(0) == b"ok"
    assert L.bvh_status_string(4) == b"non-finite input"


def test_node_layout_matches_header():
    import numpy as np
    import bvh_amd
    assert bvh_amd.NODE_DTYPE.itemsize == 48
    assert bvh_amd.NODE_DTYPE.fields["left"][1] == 32
    assert bvh_amd.NODE_DTYPE.fields["right"][1] == 36
    assert np.dtype(np.float32).itemsize == 4
