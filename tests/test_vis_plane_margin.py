"""Pins the plane-side rule of the few-occluder visibility walk (CPU only).

The GPU's walk over at most 4 occluder triangles leaves a triangle's
Moller-Trumbore test out when both end points of a ray lie beyond a margin on
one side of the triangle's plane: side(p) = N.p - N.v0 with N = e1 x e2 and
margin m = 1e-4 scale |e1|_1 |e2|_1, all in FP32 in that order (vis.hip
plane_side, plane_margin). That is only exact if the FP32 test could never
accept such a ray. These tests check the rule against the oracle's own FP32
triangle test (the same arithmetic as the GPU's, vis_oracle.h) on random rays,
rays that graze the plane just beyond the margin, and sliver triangles: no ray
whose ends share a side beyond the margin is a hit, while rays that cross the
plane are (so the check is not vacuous), and with no margin at all the rule
would be wrong (so the margin is needed, and the grazing rays find that).
"""
import numpy as np

import oracle

F32 = np.float32


def _scene(tri, far):
    s = oracle.Scene()
    s.set_bruteforce(True)
    s.set_initial_conditions(1480.0, 1e3, 0.0)
    s.load_geometry(np.asarray(tri, F32).reshape(1, 3, 3), 0, 10.0)
    s.load_source_points(np.array([[far, 0.0, 0.0]], F32))
    s.load_field_points(np.array([[0.0, far, 0.0]], F32))
    s.render()
    return s


def _dot(n, p):
    p = np.asarray(p, F32).reshape(-1, 3)
    return ((n[0] * p[:, 0] + n[1] * p[:, 1]).astype(F32) + n[2] * p[:, 2]).astype(F32)


def _plane(tri, scale):
    """N, N.v0 and the margin as the GPU forms them (FP32, same order)."""
    a, b, c = (np.asarray(v, F32) for v in tri)
    e1, e2 = (b - a).astype(F32), (c - a).astype(F32)
    n = np.array([e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]], F32)
    off = _dot(n, a)[0]
    l1 = F32(F32(abs(e1[0]) + abs(e1[1])) + abs(e1[2]))
    l2 = F32(F32(abs(e2[0]) + abs(e2[1])) + abs(e2[2]))
    m = F32(F32(F32(F32(1e-4) * F32(scale)) * l1) * l2)
    return n, off, m


def _rays(o, t):
    v = (t - o).astype(F32)
    r = np.sqrt(((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]).astype(F32) + v[:, 2] * v[:, 2]).astype(F32)).astype(F32)
    d = (v * (F32(1) / r)[:, None]).astype(F32)
    return d, r


def _cases(rng):
    tris = [rng.uniform(-1.5, 1.5, (3, 3)) for _ in range(12)]
    # slivers: one short edge, and one nearly collinear vertex
    tris.append(np.array([[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1e-3, 0.0]]))
    tris.append(np.array([[-1.0, -1.0, 0.2], [1.0, 1.0, 0.2], [0.0, 1e-3, 0.2]]))
    tris.append(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]))  # an axis plane, as the trihedral's
    return tris


def _check(margin_factor):
    """(rays with both ends beyond the margin on one side that the FP32 test accepts,
    those it rejects, accepted rays that cross the plane) over all cases."""
    rng = np.random.default_rng(11)
    far = 3.0
    bad = same_total = crossing_hits = 0
    for tri in _cases(rng):
        s = _scene(tri, far)
        scale = float(s.vis_params()[0])
        n, off, m0 = _plane(tri, scale)
        m = F32(m0 * F32(margin_factor))
        nn = n.astype(np.float64)
        nhat = nn / np.linalg.norm(nn)
        a, b, c = (np.asarray(v, np.float64) for v in tri)
        # random end points anywhere in the scene box
        o1 = rng.uniform(-scale, scale, (40000, 3))
        t1 = rng.uniform(-scale, scale, (40000, 3))
        # grazing segments: through a point of the triangle's plane near the triangle, almost parallel to
        # it (tilt 0 to 1e-5), lifted off the plane by 1-3 x 1e-7 scale / |N| (far inside the margin:
        # where the rule without a margin fails) or by 1-3 margins (just beyond it)
        k = 80000
        w1, w2 = rng.random(k), rng.random(k)
        inside = a + (w1 * (1 - w2))[:, None] * (b - a) + (w1 * w2)[:, None] * (c - a)
        u = rng.normal(size=(k, 3))
        u -= (u @ nhat)[:, None] * nhat
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        tilt = rng.choice([0.0, 1e-8, 1e-7, 1e-6, 1e-5], size=k) * rng.choice([-1, 1], size=k)
        dvec = u + tilt[:, None] * nhat
        dvec /= np.linalg.norm(dvec, axis=1, keepdims=True)
        lift = np.where(np.arange(k) < k // 2, 1e-7 * scale, float(m0)) / np.linalg.norm(nn)
        h = rng.uniform(1.0, 3.0, k) * lift * rng.choice([-1, 1], size=k)
        length = rng.uniform(0.01, 2.0 * np.sqrt(3.0) * scale * 0.95, k)
        start = inside - (0.5 * length)[:, None] * dvec + h[:, None] * nhat
        o = np.concatenate([o1, start]).astype(F32)
        t = np.concatenate([t1, start + length[:, None] * dvec]).astype(F32)
        keep = (np.abs(o) <= scale).all(1) & (np.abs(t) <= scale).all(1)
        o, t = o[keep], t[keep]
        d, r = _rays(o, t)
        ok = r > F32(2e-3)
        o, t, d, r = o[ok], t[ok], d[ok], r[ok]
        so, st = (_dot(n, o) - off).astype(F32), (_dot(n, t) - off).astype(F32)
        same = ((so > m) & (st > m)) | ((so < -m) & (st < -m))
        for tmin in (F32(0.0), F32(1e-3)):
            hit = s.anyhit(o, d, tmin, (r - F32(1e-3)).astype(F32), brute=True)
            bad += int((hit & same).sum())
            crossing_hits += int((hit & ~same).sum())
        same_total += int(same.sum())
    return bad, same_total, crossing_hits


def test_rays_beyond_the_plane_margin_never_hit():
    bad, same_total, crossing_hits = _check(1.0)
    assert bad == 0, f"{bad} accepted hits with both ends beyond the plane margin"
    assert same_total > 400000
    assert crossing_hits > 1000


def test_the_margin_is_needed():
    """With no margin (sides compared with zero) the grazing rays include accepted
    hits whose computed end-point sides agree: the rule without its margin would
    change the visibility words."""
    bad, _, _ = _check(0.0)
    assert bad > 0
