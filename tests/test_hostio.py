"""Host-side I/O (CPU): STL round trip and reader, surface colour map."""
import os

import numpy as np
import pytest

import scenes

REF_SPHERE = "/root/reference/run_scripts/sphere_1m_radius.stl"


def test_stl_round_trip(tmp_path, api):
    tris, _ = scenes.icosphere(2)
    p = tmp_path / "s.stl"
    api.write_stl(p, tris)
    assert os.path.getsize(p) == 84 + 50 * len(tris)
    back = api.read_stl(p)
    assert np.array_equal(back, tris)


def test_stl_ascii(tmp_path, api):
    txt = "solid t\n facet normal 0 0 1\n  outer loop\n   vertex 0 0 0\n   vertex 1 0 0\n   vertex 0 1 0\n" \
          "  endloop\n endfacet\nendsolid t\n"
    p = tmp_path / "a.stl"
    p.write_text(txt)
    t = api.read_stl(p)
    assert t.shape == (1, 3, 3) and np.array_equal(t[0], np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32))


def test_load_stl_creates_object(tmp_path, api):
    p = tmp_path / "p.stl"
    api.write_stl(p, scenes.make_rectangle(1, 1))
    api.load_stl(p, 0, 0.05)
    assert api.object_count(0) == 1
    with pytest.raises(api.SonarRenderError):
        api.load_stl(tmp_path / "missing.stl", 0, 0.05)


@pytest.mark.skipif(not os.path.exists(REF_SPHERE), reason="reference input mesh not mounted")
def test_reference_sphere_mesh_matches_icosphere(api):
    """The reference's sphere input (5120 triangles, 2562 vertices) is the 4-times split icosahedron."""
    ref = api.read_stl(REF_SPHERE)
    mine, _ = scenes.icosphere(4)
    assert ref.shape == (5120, 3, 3)
    a = np.unique(np.round(ref.reshape(-1, 3).astype(np.float64), 5), axis=0)
    b = np.unique(np.round(mine.reshape(-1, 3).astype(np.float64), 5), axis=0)
    assert len(a) == 2562
    r = np.linalg.norm(ref.reshape(-1, 3).astype(np.float64), axis=1)
    assert np.all(np.abs(r - 1) < 1e-6)
    # same vertex set up to a rigid rotation is not assumed; compare radii and counts, and face areas
    area = lambda t: np.sort(0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1))
    assert np.allclose(area(ref.astype(np.float64)), area(mine.astype(np.float64)), rtol=1e-5)


def test_colour_map_levels_and_wheel(api):
    # brightness: 20 log10(|p| + 1e-8) mapped from [-90, -20] dB to [0, 1], grey at saturation 0.8
    mags = np.array([1e-6, 10 ** (-55 / 20), 1.0])
    rgba = api.colour_map(mags.astype(complex), -90, -20, phase=False)
    lvl = np.clip((20 * np.log10(mags + 1e-8) + 90) / 70, 0, 1)
    assert np.allclose(rgba[:, 0], 0.8 * lvl, atol=1e-6) and np.allclose(rgba[:, 3], 1)
    # phase wheel: arg p = -pi .. pi maps to hue 0 .. 1; value = brightness, saturation 0.8
    ph = np.array([-np.pi + 1e-3, -np.pi / 2, 0.0, np.pi / 2])
    rgba = api.colour_map(np.exp(1j * ph), -90, -20, phase=True)
    assert np.allclose(rgba.max(axis=1)[:3 - 0], 1.0, atol=1e-6)  # full brightness at |p| = 1
    # hue 0 -> red dominant; hue 0.25 (arg -pi/2) -> green dominant; hue 0.5 (arg 0) -> cyan; hue 0.75 -> blue/magenta
    assert np.argmax(rgba[0, :3]) == 0
    assert np.argmax(rgba[1, :3]) == 1
    assert rgba[2, 1] > 0.99 and rgba[2, 2] > 0.99 and abs(rgba[2, 0] - 0.2) < 1e-5
    assert np.argmax(rgba[3, :3]) == 2


FIXTURE_TRIHEDRAL = os.path.join(os.path.dirname(__file__), "golden", "thick_trihedral.stl")


def test_thick_trihedral_fixture(api):
    """The plate-thick trihedral (tests/golden/make_thick_trihedral.py, the geometry of the reference's
    run_scripts/trihedral_reflector.stl): 36 triangles, three closed 1 x 1 x 0.01 m boxes with outward
    normals, total area 6.12 m^2; the STL reader returns what the generator wrote."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
    import make_thick_trihedral
    t = api.read_stl(FIXTURE_TRIHEDRAL).astype(np.float64)
    assert t.shape == (36, 3, 3)
    assert np.array_equal(t.astype(np.float32), make_thick_trihedral.thick_trihedral())
    area = 0.5 * np.linalg.norm(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]), axis=1).sum()
    assert abs(area - 6.12) < 1e-5
    assert np.allclose(np.abs(t).max(), np.sqrt(0.5), atol=1e-6)
    from collections import Counter
    for box in t.reshape(3, 12, 3, 3):
        edges = Counter()
        for tri in box:
            for a, b in ((0, 1), (1, 2), (2, 0)):
                edges[tuple(sorted([tuple(np.round(tri[a], 5)), tuple(np.round(tri[b], 5))]))] += 1
        assert set(edges.values()) == {2}  # closed
        # divergence theorem: outward normals give the box volume 1 x 1 x 0.01
        vol = sum(np.dot(tri[0], np.cross(tri[1], tri[2])) for tri in box) / 6.0
        assert abs(vol - 0.01) < 1e-6
