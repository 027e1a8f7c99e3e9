"""Writes thick_trihedral.stl: a plate-thick trihedral corner reflector built from
its geometry (the shape of run_scripts/trihedral_reflector.stl in the reference):
three mutually perpendicular 1 m x 1 m square plates, 1 cm thick, each a closed box
of 12 outward-facing triangles, the squares set diamond-wise (corners on the two
in-plane axes at +-sqrt(1/2) m) in the planes z = 0, x = 0 and y = 0, thickened
towards +z, +x and -y. 36 triangles, total area 6 + 12 x 0.01 = 6.12 m^2.

Run: python tests/golden/make_thick_trihedral.py (numpy only; binary STL).
"""
import os
import struct

import numpy as np

H = float(np.sqrt(0.5))
T = 0.01


def plate(axis, sign):
    """Box of one plate: the diamond in the plane normal to `axis`, thickness T along sign * axis."""
    u, v = [a for a in range(3) if a != axis]
    ring = []
    for pu, pv in ((0.0, -H), (H, 0.0), (0.0, H), (-H, 0.0)):  # diamond corners, counter-clockwise in (u, v)
        p = np.zeros(3)
        p[u], p[v] = pu, pv
        ring.append(p)
    lo = [p.copy() for p in ring]
    hi = [p + sign * T * np.eye(3)[axis] for p in ring]
    tris = [[lo[0], lo[2], lo[1]], [lo[0], lo[3], lo[2]],  # bottom face
            [hi[0], hi[1], hi[2]], [hi[0], hi[2], hi[3]]]  # top face
    for k in range(4):  # side faces
        a, b = k, (k + 1) % 4
        tris += [[lo[a], lo[b], hi[b]], [lo[a], hi[b], hi[a]]]
    tris = np.array(tris)
    # orient every triangle away from the box centre
    c = (sum(lo) + sum(hi)) / 8.0
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(axis=1) - c) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris


def thick_trihedral():
    return np.concatenate([plate(2, +1.0), plate(0, +1.0), plate(1, -1.0)]).astype(np.float32)


def write_stl(path, tris):
    with open(path, "wb") as f:
        f.write(b"thick trihedral corner reflector (make_thick_trihedral.py)".ljust(80, b" "))
        f.write(struct.pack("<I", len(tris)))
        for t in tris.astype(np.float64):
            n = np.cross(t[1] - t[0], t[2] - t[0])
            n = n / np.linalg.norm(n)
            f.write(struct.pack("<12fH", *n, *t.reshape(9), 0))


if __name__ == "__main__":
    write_stl(os.path.join(os.path.dirname(os.path.abspath(__file__)), "thick_trihedral.stl"), thick_trihedral())
