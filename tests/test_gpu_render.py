"""The renderer (renderImage of the reference API, through bvh_render_ppm):
every shading mode's PPM bytes equal the reference's shading definitions
(render_engine.cu:362-446 kernel, :519-593 heatmap post-process, :486-510
PPM writer) evaluated in float32 on the oracle's closest hits for the same
host-generated camera rays."""
import numpy as np
import pytest

import bvh_amd as B
import oracle as O

pytestmark = pytest.mark.gpu

F = np.float32


def _norm(v):
    ln = np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])
    small = ln < F(1e-12)
    inv = F(1.0) / np.where(small, F(1.0), ln)
    out = np.stack([v[0] * inv, v[1] * inv, v[2] * inv])
    out[:, small] = F(0.0)
    return out


def _turbo_table():
    import re
    import os
    src = open(os.path.join(os.path.dirname(B.__file__), "host", "render.cpp")).read()
    body = src[src.index("kTurbo1e5[256 * 3] = {") + 22:]
    body = body[:body.index("};")]
    ints = np.array([int(x) for x in re.findall(r"\d+", body)], np.int64)
    assert ints.size == 768
    return (ints.astype(F) / F(100000)).reshape(256, 3)


def _expected(nodes, mesh, rays, mode, hmax_fixed=0.0):
    t, prim, cnt = O.trace(nodes, mesh.soa, rays)
    npx = rays.shape[0]
    rgb = np.empty((npx, 3), F)
    rgb[:] = np.array([0.15, 0.15, 0.18], F)
    hit = prim >= 0
    s = mesh.soa
    p = prim[hit]
    e1 = np.stack([s[3][p] - s[0][p], s[4][p] - s[1][p], s[5][p] - s[2][p]])
    e2 = np.stack([s[6][p] - s[0][p], s[7][p] - s[1][p], s[8][p] - s[2][p]])
    n = np.stack([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]])
    n = _norm(n)
    if mode == 0:
        rgb[hit] = ((n + F(1.0)) * F(0.5)).T
    elif mode == 1:
        lo, hi = nodes["bmin"][0], nodes["bmax"][0]
        e = hi - lo
        diag = np.sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2])
        diag = F(1.0) if diag < F(1e-6) else diag
        depth = np.clip(t[hit] / diag, F(0.0), F(1.0))
        rgb[hit] = (F(1.0) - depth)[:, None]
    elif mode == 2:
        light = _norm(np.array([[0.5], [0.8], [0.6]], F))[:, 0]
        d = rays[hit, 3:].T
        flip = (n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) > F(0.0)
        n = np.where(flip, -n, n)
        diff = np.maximum(F(0.0), n[0] * light[0] + n[1] * light[1] + n[2] * light[2])
        shade = np.minimum(F(1.0), F(0.15) + diff * F(0.85))
        rgb[hit] = np.stack([shade * (F(0.7) + F(0.3) * np.abs(n[k])) for k in range(3)]).T
    else:
        visits = cnt[:, 0]
        hmax = F(hmax_fixed)
        if hmax <= 0:
            v = np.sort(visits[visits > 0])
            k = min(v.size - 1, int(v.size * 0.999))
            hmax = max(F(1.0), F(v[k]))
        tab = _turbo_table()
        on = visits > 0
        tt = np.clip(visits[on].astype(F) / hmax, F(0.0), F(1.0))
        sc = tt * F(255.0)
        i0 = sc.astype(np.int32)
        i1 = np.minimum(i0 + 1, 255)
        fr = sc - i0.astype(F)
        rgb[on] = (tab[i0] * (F(1.0) - fr)[:, None] + tab[i1] * fr[:, None])
    v = np.clip(rgb, F(0.0), F(1.0))
    return (v * F(255.0) + F(0.5)).astype(np.uint8).reshape(-1)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_render_modes_match_reference_shading(ctx, tmp_path, mode):
    mesh = B.Mesh.blobs(17, 30_000)
    dm = ctx.upload(mesh)
    tree = ctx.tree()
    try:
        ctx.build(dm, tree, algo=B.ALGO_LBVH)
        nodes, _ = tree.download()
    finally:
        tree.free()
        dm.free()
    w, h = 96, 64
    eye, at, up, fov = (1.6, 1.1, 1.8), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 55.0
    path = str(tmp_path / f"m{mode}.ppm")
    st = B.render_ppm(nodes, mesh, w, h, eye, at, up, fov, mode=mode, path=path)
    raw = open(path, "rb").read()
    header = f"P6\n{w} {h}\n255\n".encode()
    assert raw[:len(header)] == header
    got = np.frombuffer(raw[len(header):], np.uint8)
    rays = B.rays_camera(eye, at, up, fov, w, h)
    want = _expected(nodes, mesh, rays, mode)
    np.testing.assert_array_equal(got, want)
    assert st["avg_nodes_visited"] > 0


def test_turbo_table_endpoints_and_monotone_lightness():
    """Published Turbo endpoints (dark blue at 0, dark red at 1) and the table
    size; the values come from the C++ source the renderer compiles."""
    tab = _turbo_table()
    np.testing.assert_array_equal(tab[0], np.array([0.18995, 0.07176, 0.23217], F))
    np.testing.assert_array_equal(tab[255], np.array([0.47960, 0.01583, 0.01055], F))
