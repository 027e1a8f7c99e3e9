"""Scene builders (numpy only) for the workloads the engine is measured on.

Mesh conventions follow the reference's helpers (run_scripts/modules/geo.py):
``make_rectangle`` (geo.py:75-113), ``generate_field_points`` (geo.py:64-73,
x = R sin(theta), z = R cos(theta)) and ``halve_facets`` (geo.py:158-196). A
facet's normal is unit((v1-v2) x (v1-v3)) (src/Facet.cpp:102-105), so every
builder orders vertices counter-clockwise seen from the lit side.

Every mesh is returned as a float32 array of shape (N, 3, 3).
"""
import math

import numpy as np

C_WATER = 1480.0


def wavelength(f, c=C_WATER):
    return c / f


def make_rectangle(length, width, xy_plane=True):
    """Two-triangle plate centred on the origin (geo.py:75-113)."""
    if xy_plane:
        v = np.array([[-length / 2, -width / 2, 0.0], [length / 2, -width / 2, 0.0],
                      [length / 2, width / 2, 0.0], [-length / 2, width / 2, 0.0]])
    else:
        v = np.array([[-length / 2, 0.0, -width / 2], [length / 2, 0.0, -width / 2],
                      [length / 2, 0.0, width / 2], [-length / 2, 0.0, width / 2]])
    return np.array([[v[0], v[1], v[2]], [v[0], v[2], v[3]]], dtype=np.float32)


def generate_field_points(radius, angles_deg):
    """Points on an arc in the x-z plane (geo.py:64-73)."""
    a = np.radians(np.asarray(angles_deg, dtype=np.float64))
    return np.stack([radius * np.sin(a), np.zeros_like(a), radius * np.cos(a)], axis=1).astype(np.float32)


def halve_facets(tris):
    """Split each triangle into four at its edge midpoints (geo.py:158-196)."""
    t = np.asarray(tris, dtype=np.float32)
    v1, v2, v3 = t[:, 0], t[:, 1], t[:, 2]
    m12, m23, m31 = (v1 + v2) / 2, (v2 + v3) / 2, (v3 + v1) / 2
    out = np.stack([np.stack([v1, m12, m31], 1), np.stack([m12, v2, m23], 1),
                    np.stack([m23, v3, m31], 1), np.stack([m12, m23, m31], 1)], 1)
    return out.reshape(-1, 3, 3)


def icosphere(subdivisions=4, radius=1.0):
    """Geodesic sphere: icosahedron split `subdivisions` times, vertices on the sphere.

    subdivisions=4 gives 5120 triangles and 2562 vertices, the mesh of the
    reference's run_scripts/sphere_1m_radius.stl. Outward normals.
    """
    p = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p),
             (0, -1, -p), (0, 1, -p), (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius
    F = np.array(faces)
    tris = V[F]
    # orient outward: (v2-v1) x (v3-v1) . centroid > 0
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(axis=1)) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris.astype(np.float32), len(V)


def trihedral(leg=1.0):
    """Three mutually perpendicular right-isosceles plates, normals into the octant."""
    o = [0.0, 0.0, 0.0]
    return np.array([[o, [leg, 0, 0], [0, leg, 0]],
                     [o, [0, leg, 0], [0, 0, leg]],
                     [o, [0, 0, leg], [leg, 0, 0]]], dtype=np.float32)


def octant_incidences(n_side=8, radius=15.0, seed=0):
    """Seeded jittered n_side x n_side (theta, phi) grid over the octant interior."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(n_side), np.arange(n_side), indexing="ij")
    th = (i + rng.uniform(0.05, 0.95, i.shape)) / n_side * (np.pi / 2)
    ph = (j + rng.uniform(0.05, 0.95, j.shape)) / n_side * (np.pi / 2)
    th, ph = th.ravel(), ph.ravel()
    return (radius * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)).astype(np.float32)


# ---------------------------------------------------------------------------
# Workload definitions. Each returns a dict consumed by run_scene() / tests.

def config_plate(f=10e3, c=C_WATER, frac=7.5, n_points=181):
    """1x1 m plate, monopole 5 m away 20 deg off normal, 181-point 10 m arc."""
    lam = c / f
    a = math.radians(20.0)
    return dict(
        name="plate_1m_10kHz", f=f, c=c, atten=0.0,
        targets=[(make_rectangle(1.0, 1.0), lam / frac)], fields=[],
        sources=np.array([[5 * math.sin(a), 0.0, 5 * math.cos(a)]], np.float32),
        points=generate_field_points(10.0, np.linspace(-90, 90, n_points)),
        bounces=0, project_field_objects=False)


def config_sphere(f=10e3, c=C_WATER, frac=7.5, subdiv=4, n_points=360, surface_cells=256, field_surface=True):
    """Rigid r=1 m icosphere, monopole 20 m away, 360-point 30 m ring, field plane."""
    lam = c / f
    tris, _ = icosphere(subdiv)
    dfld = lam / frac
    fields = []
    if field_surface:
        L = surface_cells * dfld
        fields = [(make_rectangle(L, L, xy_plane=False), dfld)]
    return dict(
        name="sphere_r1_10kHz", f=f, c=c, atten=0.0,
        targets=[(tris, lam / frac)], fields=fields,
        sources=np.array([[0.0, 0.0, 20.0]], np.float32),
        points=generate_field_points(30.0, np.arange(n_points) * (360.0 / n_points)),
        bounces=0, project_field_objects=False)


def config_piston(f=10e3, c=C_WATER, frac=7.5, n_points=91, field_surface=False):
    """A 0.2 m square piston source object 2 m above a 1 m plate, 20 degrees off
    its normal, facing it; field points on a 10 m arc; optional field plane."""
    lam = c / f
    a = math.radians(20.0)
    # piston: a small plate rotated to face the origin from (2 sin a, 0, 2 cos a)
    pis = make_rectangle(0.2, 0.2)[:, :, :] * np.float32(1.0)
    pis = pis[:, [0, 2, 1]]                       # reversed winding: normal -z
    rot = np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])
    centre = np.array([2 * math.sin(a), 0.0, 2 * math.cos(a)])
    pis = (pis.reshape(-1, 3) @ rot.T + centre).reshape(-1, 3, 3).astype(np.float32)
    fields = []
    if field_surface:
        fields = [(make_rectangle(1.5, 1.5, xy_plane=False) + np.float32([0.0, 0.0, 1.0]), lam / frac)]
    return dict(
        name="piston_plate_10kHz", f=f, c=c, atten=0.0,
        targets=[(make_rectangle(1.0, 1.0), lam / frac)], fields=fields,
        source_objects=[(pis, lam / frac)],
        sources=np.zeros((0, 3), np.float32),
        points=generate_field_points(10.0, np.linspace(-90, 90, n_points)),
        bounces=0, project_field_objects=field_surface)


def config_trihedral(f=20e3, c=C_WATER, frac=7.5, n_side=8, bounces=3, radius=15.0):
    """Trihedral corner reflector, 1 m legs, batched octant incidences, monostatic."""
    lam = c / f
    src = octant_incidences(n_side, radius)
    return dict(
        name="trihedral_1m_20kHz", f=f, c=c, atten=0.0,
        targets=[(trihedral(1.0), lam / frac)], fields=[],
        sources=src, points=src.copy(), bounces=bounces, project_field_objects=False,
        batched=True)


def _orient_outward(tris, inside):
    """Reorders each triangle's vertices so that unit((v1-v2) x (v1-v3)) points away from inside(centroid)."""
    t = np.asarray(tris, dtype=np.float64).copy()
    n = np.cross(t[:, 0] - t[:, 1], t[:, 0] - t[:, 2])
    c = t.mean(axis=1)
    flip = np.einsum("ij,ij->i", n, c - inside(c)) < 0
    t[flip] = t[flip][:, [0, 2, 1]]
    return t


def capped_cylinder(length=6.0, radius=0.5, n_around=48, n_along=24, n_cap=8):
    """Cylinder along x (body from -length/2 to +length/2) with hemispherical end caps; outward normals."""
    half = length / 2.0
    phi = np.linspace(0.0, 2.0 * np.pi, n_around + 1)
    xs = np.linspace(-half, half, n_along + 1)
    tris = []
    for i in range(n_along):
        for j in range(n_around):
            a = (xs[i], radius * np.cos(phi[j]), radius * np.sin(phi[j]))
            b = (xs[i + 1], radius * np.cos(phi[j]), radius * np.sin(phi[j]))
            c = (xs[i + 1], radius * np.cos(phi[j + 1]), radius * np.sin(phi[j + 1]))
            d = (xs[i], radius * np.cos(phi[j + 1]), radius * np.sin(phi[j + 1]))
            tris += [(a, b, c), (a, c, d)]
    th = np.linspace(0.0, np.pi / 2, n_cap + 1)  # 0 at the rim, pi/2 at the pole
    for sgn in (-1.0, 1.0):
        def p(k, j):
            return (sgn * (half + radius * np.sin(th[k])), radius * np.cos(th[k]) * np.cos(phi[j]),
                    radius * np.cos(th[k]) * np.sin(phi[j]))
        for k in range(n_cap):
            for j in range(n_around):
                if k == n_cap - 1:
                    tris.append((p(k, j), p(k, j + 1), p(k + 1, 0)))
                else:
                    tris += [(p(k, j), p(k, j + 1), p(k + 1, j + 1)), (p(k, j), p(k + 1, j + 1), p(k + 1, j))]
    t = np.array(tris, dtype=np.float64)

    def inside(c):  # nearest point of the axis segment
        return np.stack([np.clip(c[:, 0], -half, half), np.zeros(len(c)), np.zeros(len(c))], 1)
    return _orient_outward(t, inside).astype(np.float32)


def box(lo, hi):
    """Axis-aligned closed box, 12 triangles, outward normals."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    v = np.array([[lo[0] if i & 1 == 0 else hi[0], lo[1] if i & 2 == 0 else hi[1], lo[2] if i & 4 == 0 else hi[2]]
                  for i in range(8)])
    quads = [(0, 1, 3, 2), (4, 5, 7, 6), (0, 1, 5, 4), (2, 3, 7, 6), (0, 2, 6, 4), (1, 3, 7, 5)]
    t = np.array([(v[a], v[b], v[c]) for a, b, c, d in quads] + [(v[a], v[c], v[d]) for a, b, c, d in quads])
    centre = (lo + hi) / 2
    return _orient_outward(t, lambda c: np.broadcast_to(centre, c.shape)).astype(np.float32)


def hull_with_fin(length=6.0, radius=0.5, n_around=48, n_along=24, n_cap=8):
    """Capped-cylinder hull with a 1 x 0.3 x 0.05 m box fin on top (its foot 5 cm inside the hull)."""
    fin = box((-0.5, -0.025, radius - 0.05), (0.5, 0.025, radius + 0.25))
    return np.concatenate([capped_cylinder(length, radius, n_around, n_along, n_cap), fin])


def config_hull(f=8e3, c=C_WATER, frac=7.5, n_angles=720, radius=50.0, atten_db_per_m=0.5, bounces=2,
                n_around=48, n_along=24, n_cap=8, field_cells=0):
    """Capped-cylinder hull with a fin, monostatic angles around it in the horizontal plane.

    Attenuation is given in dB/m and converted to the Np/m the pair kernel takes
    (the reference feeds the dB/m figure straight into Im k, mh.py:72; here the
    physical conversion is used). field_cells > 0 adds a field-surface plane of
    that many cells a side (x-y plane, 2 m below the axis).
    """
    lam = c / f
    a = np.arange(n_angles) * (2.0 * np.pi / n_angles)
    src = (radius * np.stack([np.cos(a), np.sin(a), np.zeros_like(a)], 1)).astype(np.float32)
    fields = []
    if field_cells:
        d = lam / frac
        L = field_cells * d
        plate = make_rectangle(L, L)
        plate[:, :, 2] -= 2.0
        fields.append((plate, d))
    return dict(name=f"hull_6m_{f / 1e3:g}kHz", f=f, c=c, atten=atten_db_per_m / (20.0 / math.log(10.0)),
                targets=[(hull_with_fin(6.0, 0.5, n_around, n_along, n_cap), lam / frac)], fields=fields,
                sources=src, points=src.copy(), bounces=bounces, project_field_objects=bool(fields),
                batched=True)


def random_convex_polyhedron(rng, ntri_min=8, ntri_max=40, extent=(0.3, 1.5)):
    """Convex hull of random points: 8-40 outward-facing triangles, seeded."""
    from scipy.spatial import ConvexHull
    while True:
        npts = int(rng.integers(6, 22))
        pts = rng.normal(size=(npts, 3))
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
        pts *= rng.uniform(0.6, 1.0, (npts, 1))
        hull = ConvexHull(pts)
        if ntri_min <= len(hull.simplices) <= ntri_max:
            break
    size = rng.uniform(*extent)
    span = (pts.max(0) - pts.min(0)).max()
    pts = pts * (size / span)
    tris = pts[hull.simplices]
    c = pts.mean(0)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(1) - c) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris.astype(np.float32)


def config_batch(n_scenes=256, seed=0, frac=4.0, n_points=32, bounces=2, c=C_WATER):
    """Batched uneven scenes: convex polyhedra or plates, f in 2-20 kHz, fragments lambda/frac."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        f = float(rng.uniform(2e3, 20e3))
        if rng.random() < 0.25:
            a, b = rng.uniform(0.3, 1.5, 2)
            tris = make_rectangle(a, b)
        else:
            tris = random_convex_polyhedron(rng)
        u = rng.normal(size=3)
        u[2] = abs(u[2]) + 0.3
        u /= np.linalg.norm(u)
        src = (10.0 * u)[None].astype(np.float32)
        ang = np.sort(rng.uniform(-90, 90, n_points))
        pts = generate_field_points(10.0, ang)
        out.append(dict(name=f"batch_{s}", f=f, c=c, atten=0.0, targets=[(tris, c / f / frac)], fields=[],
                        sources=src, points=pts, bounces=bounces))
    return out


def scene_pairs(scene, n_act, n_cells_field=None):
    """Ordered pair count of one scene run (the Gpairs/s unit of work).

    One pair is one ordered (emitter fragment with A > 1e-6, receiver)
    contribution counted before culling; reciprocity counts 2 per unordered
    pair; sources and field points count as single emitters/receivers.
    ``n_act`` = per-target list of (active fragments, cells, per-facet active counts);
    ``n_cells_field`` = active cells of each field object.
    """
    ns = len(scene["sources"])
    nq = len(scene["points"])
    planes = ns if scene.get("batched") else 1
    per_plane_src = 1 if scene.get("batched") else ns
    total = 0
    for nact, ncell, per_facet in n_act:
        total += per_plane_src * nact  # point->fragment
        self_pairs = nact * nact - sum(k * k for k in per_facet)
        total += scene["bounces"] * self_pairs
        total += nact * nq
        if n_cells_field:
            total += nact * sum(n_cells_field)
    return total * planes
