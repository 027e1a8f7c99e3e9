"""Host render (RenderOpenGL's picture without OpenGL), CPU only.

Checks the camera (glm::lookAt + 45-degree perspective, src/ModelGl.cpp:94-104),
the per-facet texture mapping (Facet.cpp:20-65: cell (i, j) of a facet is
texel (i, j), the base edge at v = 0), the colour map of MakeSurface
(FacetCuda.cu:43-72), the depth test, the beige background and the PNG file.
"""
import math
import struct
import zlib

import numpy as np
import pytest

BEIGE = (245, 245, 219)


def _grey(db, min_db=-90.0, max_db=-20.0):
    mag = min(max((db - min_db) / (max_db - min_db), 0.0), 1.0)
    return round(0.8 * mag * 255)


def _cells(api, tris, delta):
    grids = [api.fragment_facet(t, delta) for t in tris]
    return grids, sum(g[0][0] * g[0][1] for g in grids)


def _read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n, typ = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert zlib.crc32(typ + body) & 0xffffffff == crc
        if typ == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert (depth, ctype) == (8, 2)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()
    return rows[:, 1:].reshape(h, w, 3)


def test_uniform_plate_colour_and_background():
    import scenes
    import sonar_api as api
    tris = scenes.make_rectangle(1.0, 1.0)
    delta = 0.05
    _, n = _cells(api, tris, delta)
    p = np.full(n, 10 ** (-50 / 20), complex)  # -50 dB everywhere
    img = api.render_facets_rgb(tris, delta, p, 64, 48, [0, 0, 3, 0, 0, 0, 0, 1, 0])
    g = _grey(-50.0)
    assert tuple(img[24, 32]) == (g, g, g)
    assert tuple(img[0, 0]) == BEIGE and tuple(img[47, 63]) == BEIGE
    # the plate's footprint: half-width 0.5 at distance 3 under a 45-degree vertical field of view
    half = 0.5 / (3 * math.tan(math.radians(22.5))) * 24  # pixels from the centre row
    col = img[:, 32, 0]
    inside = np.where(col == g)[0]
    assert abs((inside.max() - inside.min() + 1) / 2 - half) <= 1.0


def test_orientation_and_texture_axes():
    """World +x is image right and +y image up for eye +z, up +y; the texture
    follows the facet's own cell grid (a gradient along the cells' x shows along
    the facet's base edge)."""
    import scenes
    import sonar_api as api
    tris = scenes.make_rectangle(1.0, 1.0)
    delta = 0.1
    grids, n = _cells(api, tris, delta)
    # per cell: loud where the cell's emitter point has x > 0.2, quiet elsewhere
    pts = np.concatenate([g[2] for g in grids])
    p = np.where(pts[:, 0] > 0.2, 10 ** (-25 / 20), 10 ** (-85 / 20)).astype(complex)
    img = api.render_facets_rgb(tris, delta, p, 80, 80, [0, 0, 2.5, 0, 0, 0, 0, 1, 0])
    row = img[40, :, 0].astype(int)
    right, left = row[52], row[28]
    assert right > 150 and left < 30, (left, right)
    # looking from -z instead mirrors the picture
    img2 = api.render_facets_rgb(tris, delta, p, 80, 80, [0, 0, -2.5, 0, 0, 0, 0, 1, 0])
    assert img2[40, 28, 0] > 150 and img2[40, 52, 0] < 30


def test_depth_test_nearest_wins_in_either_order():
    import scenes
    import sonar_api as api
    near = scenes.make_rectangle(0.4, 0.4) + np.float32([0, 0, 0.5])
    far = scenes.make_rectangle(1.0, 1.0)
    delta = 0.05
    _, nn = _cells(api, near, delta)
    _, nf = _cells(api, far, delta)
    view = [0, 0, 3, 0, 0, 0, 0, 1, 0]
    for order in (0, 1):
        tris = np.concatenate([near, far] if order == 0 else [far, near])
        pn = np.full(nn, 10 ** (-30 / 20), complex)
        pf = np.full(nf, 10 ** (-70 / 20), complex)
        p = np.concatenate([pn, pf] if order == 0 else [pf, pn])
        img = api.render_facets_rgb(tris, delta, p, 64, 64, view)
        assert img[32, 32, 0] == _grey(-30.0)     # the near plate covers the centre
        assert img[32, 22, 0] == _grey(-70.0)     # the far plate shows beside it


def test_phase_colours():
    import scenes
    import sonar_api as api
    tris = scenes.make_rectangle(1.0, 1.0)
    delta = 0.05
    _, n = _cells(api, tris, delta)
    amp = 10 ** (-20 / 20)  # full brightness
    # HSV with s = 0.8, v = 1: chroma 0.8, offset 0.2. hue (arg + pi) / 2 pi:
    # arg ~ pi -> hue ~ 1 (red sector end: (1, 0.2, 0.2)); arg -pi/2 -> hue 1/4 ((0.6, 1, 0.2))
    for ph, rgb in [(math.pi - 1e-4, (255, 51, 51)), (-math.pi / 2, (153, 255, 51))]:
        p = np.full(n, amp * np.exp(1j * ph))
        img = api.render_facets_rgb(tris, delta, p, 32, 32, [0, 0, 3, 0, 0, 0, 0, 1, 0], phase=True)
        got = tuple(int(v) for v in img[16, 16])
        assert max(abs(a - b) for a, b in zip(got, rgb)) <= 1, (ph, got, rgb)


def test_png_round_trip(tmp_path):
    import sonar_api as api
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (37, 53, 3), dtype=np.uint8)
    path = str(tmp_path / "x.png")
    api.write_png(path, img)
    assert np.array_equal(_read_png(path), img)


def test_bad_camera_rejected():
    import scenes
    import sonar_api as api
    tris = scenes.make_rectangle(1.0, 1.0)
    _, n = _cells(api, tris, 0.1)
    with pytest.raises(api.SonarRenderError):
        api.render_facets_rgb(tris, 0.1, np.ones(n), 8, 8, [0, 0, 3, 0, 0, 3, 0, 1, 0])
    with pytest.raises(api.SonarRenderError):
        api.render_facets_rgb(tris, 0.1, np.ones(n), 8, 8, [0, 0, 3, 0, 0, 0, 0, 0, 1])
