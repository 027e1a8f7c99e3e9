"""Fragmenter: the library's host fragmenter against the oracle's, and both
against geometry (CPU only; the library's fragmenter runs on the host).

Pins: cell areas sum to the facet area; every cell area equals an exact
polygon-clipping area computed independently (Sutherland-Hodgman in FP64);
library == oracle bit for bit (areas and both point forms).
"""
import numpy as np
import pytest

import oracle
import scenes


def _oracle_cells(tri, delta):
    s = oracle.Scene()
    s.set_initial_conditions(1480.0, 1e4)
    s.load_geometry(np.asarray(tri, np.float32)[None], 0, delta)
    s.render()
    g = s.grids(0, 0)[0]
    c = s.cells(0, 0)
    fr = s.frames(0, 0)[0]
    s.close()
    return g, c, fr


def _random_tris(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        t = rng.uniform(-1.5, 1.5, (3, 3)).astype(np.float32)
        a = np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) / 2
        if a > 0.05:
            out.append(t)
    return out


def _clip_area(poly, x0, x1, y0, y1):
    """Area of a convex polygon clipped to [x0,x1] x [y0,y1] (Sutherland-Hodgman, FP64)."""
    def clip(pts, inside, inter):
        out = []
        for i in range(len(pts)):
            a, b = pts[i], pts[(i + 1) % len(pts)]
            ia, ib = inside(a), inside(b)
            if ia:
                out.append(a)
            if ia != ib:
                out.append(inter(a, b))
        return out

    def ix(a, b, x):
        t = (x - a[0]) / (b[0] - a[0])
        return (x, a[1] + t * (b[1] - a[1]))

    def iy(a, b, y):
        t = (y - a[1]) / (b[1] - a[1])
        return (a[0] + t * (b[0] - a[0]), y)

    pts = list(poly)
    for inside, inter in [(lambda p: p[0] >= x0, lambda a, b: ix(a, b, x0)),
                          (lambda p: p[0] <= x1, lambda a, b: ix(a, b, x1)),
                          (lambda p: p[1] >= y0, lambda a, b: iy(a, b, y0)),
                          (lambda p: p[1] <= y1, lambda a, b: iy(a, b, y1))]:
        if not pts:
            return 0.0
        pts = clip(pts, inside, inter)
    if len(pts) < 3:
        return 0.0
    x = np.array([p[0] for p in pts])
    y = np.array([p[1] for p in pts])
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


CASES = [(t, d) for t in _random_tris(12, 1) for d in (0.07, 0.13)] + \
        [(scenes.make_rectangle(1, 1)[0], 1480 / 10e3 / 7.5), (scenes.trihedral()[1], 1480 / 20e3 / 7.5)]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_library_matches_oracle_bit_exact(case):
    import sonar_api
    tri, delta = CASES[case]
    g, c, _ = _oracle_cells(tri, delta)
    (nx, ny, nneg), area, pe, pr = sonar_api.fragment_facet(tri, delta)
    assert (nx, ny, nneg) == tuple(int(v) for v in g)
    assert np.array_equal(area.ravel().view(np.uint32), c["area"].view(np.uint32))
    assert np.array_equal(pe.view(np.uint32), c["pe"].view(np.uint32))
    assert np.array_equal(pr.view(np.uint32), c["pr"].view(np.uint32))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_area_sum_and_exact_clipping(case):
    tri, delta = CASES[case]
    g, c, fr = _oracle_cells(tri, delta)
    nx, ny, nneg = (int(v) for v in g)
    area = c["area"].reshape(ny, nx).astype(np.float64)
    facet_area = 0.5 * np.linalg.norm(np.cross((tri[1] - tri[0]).astype(np.float64), (tri[2] - tri[0]).astype(np.float64)))
    assert abs(area.sum() - facet_area) <= 2e-5 * facet_area + 1e-7
    assert area.min() >= -1e-9 and area.max() <= delta * delta * (1 + 1e-5)
    # exact areas in the facet frame: base from (-lenNeg, 0) to (lenPos, 0), apex (0, H)
    lneg, lpos, h = float(fr[13]), float(fr[14]), float(fr[15])
    poly = [(-lneg, 0.0), (lpos, 0.0), (0.0, h)]
    exact = np.zeros_like(area)
    for j in range(ny):
        for i in range(nx):
            x0 = delta * (i - nneg)
            exact[j, i] = _clip_area(poly, x0, x0 + delta, delta * j, delta * (j + 1))
    err = np.abs(area - exact).max()
    assert err <= 2e-4 * delta * delta, err


def test_config_fragment_counts():
    """Fragment counts of the measured workloads (plate 10 kHz, trihedral 20 kHz)."""
    import runner
    per, cells = runner.facet_active_counts(scenes.make_rectangle(1, 1), 1480 / 10e3 / 7.5)
    assert (sum(per), cells) == (2664, 5184)
    per, cells = runner.facet_active_counts(scenes.trihedral(), 1480 / 20e3 / 7.5)
    assert sum(per) == 15768 and cells == 31104


def test_icosphere_mesh():
    tris, nv = scenes.icosphere(4)
    assert tris.shape == (5120, 3, 3) and nv == 2562
    r = np.linalg.norm(tris.reshape(-1, 3).astype(np.float64), axis=1)
    assert np.all(np.abs(r - 1) < 1e-6)
    n = np.cross(tris[:, 0] - tris[:, 1], tris[:, 0] - tris[:, 2])
    assert np.all(np.einsum("ij,ij->i", n, tris.mean(1)) > 0)  # outward under the reference's normal rule
