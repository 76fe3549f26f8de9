"""OBJ ingestion (include/bvh_obj.h) against hand-derived expectations of the
reference loader's rules (src/loader.cpp:10-74): std::istream float
extraction, std::stoi face indices, relative indices, fan triangulation,
skipped out-of-range triangles; multi-threaded parsing equals single-threaded
parsing; a face token without an index is an error.  CPU only."""
import os

import numpy as np
import pytest

import bvh_amd as B

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "obj")
FMAX = float(np.finfo(np.float32).max)

# Vertices of tests/golden/obj/rules.obj in file order, as the stream reads them.
RULES_VERTS = [
    (0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
    (1, 0, 0),            # "1 abc 3": 'a' fails the second read, the rest of the line stays 0
    (1.5, 0.25, 2),       # "1.5.25 2": the second '.' ends the first number
    (FMAX, 0, 0),         # "1e39": overflow reads as FLT_MAX and fails the stream
    (-0.0, 0.5, 5.0),     # "-0 .5 5."
    (0.001, 200, -3),     # "1e-3 2E+2 -3e0"
    (0, 0, 0),            # "1e": an exponent without digits is not a number
    (0, 0, 0),            # "0x10": reads 0, then "x10" fails
    (7, -8, 9),           # "+7 -8 9"
    (7.5, 1, 1),          # "007.5"
    (2, 2, 2),            # "2 2 2 extra"
    (3, 3, 3),            # CRLF line
]
# Triangles (1-based vertex numbers): f 1 2 3; quad 1 2 3 4 -> (1 2 3)(1 3 4);
# f -4 -3 -1 with 4 vertices read -> (1 2 4); pentagon 5..9 -> three fans;
# f 1 2 99 and f 0 1 2 skipped; f 10//1 11//1 12//1; f 13 14 has no
# triangle; the last face's -1 -2 -3 with 15 vertices read -> (15 14 13).
RULES_TRIS = [(1, 2, 3), (1, 2, 3), (1, 3, 4), (1, 2, 4), (5, 6, 7), (5, 7, 8), (5, 8, 9), (10, 11, 12), (15, 14, 13)]


def _expected_soa(verts, tris):
    v = np.array(verts, np.float32)
    soa = np.empty((9, len(tris)), np.float32)
    for t, tri in enumerate(tris):
        for k, vid in enumerate(tri):
            soa[3 * k:3 * k + 3, t] = v[vid - 1]
    return soa


@pytest.mark.parametrize("threads", [1, 4])
def test_rules_fixture(threads):
    m = B.Mesh.from_obj(os.path.join(GOLD, "rules.obj"), threads)
    want = _expected_soa(RULES_VERTS, RULES_TRIS)
    assert m.n == len(RULES_TRIS)
    np.testing.assert_array_equal(m.soa.view(np.uint32), want.view(np.uint32))   # bits (signed zero included)


def test_more_float_forms(tmp_path):
    p = tmp_path / "f.obj"
    p.write_text("v 1\nv 1e+\nv -.e5 1 1\nv 1.e5 2 3\nv 1e-50 0.000000000000000000000000000000000000001 1\n"
                 "f 1 2 3 4 5\n")
    m = B.Mesh.from_obj(str(p))
    verts = [(1, 0, 0), (0, 0, 0), (0, 0, 0), (1e5, 2, 3), (0, np.float32(1e-39), 1)]
    tris = [(1, 2, 3), (1, 3, 4), (1, 4, 5)]
    np.testing.assert_array_equal(m.soa.view(np.uint32), _expected_soa(verts, tris).view(np.uint32))


def test_face_token_without_index_is_an_error():
    with pytest.raises(B.NativeError) as e:
        B.Mesh.from_obj(os.path.join(GOLD, "bad_token.obj"))
    assert e.value.status == 1 and "line 1" in str(e.value)
    with pytest.raises(B.NativeError):
        B.Mesh.from_obj(os.path.join(GOLD, "missing.obj"))


def _write_obj(path, mesh, quads_every=5):
    """A mesh written with %.9g (exact float round trip): every triangle's three
    vertices, faces alternating absolute, relative and v/vt/vn forms."""
    lines = []
    nv = 0
    s = mesh.soa
    for t in range(mesh.n):
        for k in range(3):
            lines.append("v %.9g %.9g %.9g" % (s[3 * k, t], s[3 * k + 1, t], s[3 * k + 2, t]))
        nv += 3
        form = t % 3
        if form == 0:
            lines.append(f"f {nv - 2} {nv - 1} {nv}")
        elif form == 1:
            lines.append("f -3 -2 -1")
        else:
            lines.append(f"f {nv - 2}/1/1 {nv - 1}/1/1 {nv}//1")
        if t % 97 == 0:
            lines.append("# comment")
    path.write_text("\n".join(lines) + "\n")


def test_threads_agree_and_round_trip(tmp_path):
    m = B.Mesh.uniform(77, 30_000)
    p = tmp_path / "big.obj"
    _write_obj(p, m)
    assert p.stat().st_size > (1 << 20)   # large enough to be split across threads
    a = B.Mesh.from_obj(str(p), 1)
    b = B.Mesh.from_obj(str(p), 8)
    np.testing.assert_array_equal(a.soa.view(np.uint32), m.soa.view(np.uint32))
    np.testing.assert_array_equal(b.soa.view(np.uint32), m.soa.view(np.uint32))
