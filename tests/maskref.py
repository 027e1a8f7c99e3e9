"""Oracle-side visibility masks (test infrastructure).

A pair contributes when both cosine culls pass and the shadow ray is clear.
The culls are evaluated here in numpy FP32 in the reference's operation order
(r = sqrtf(dot), uv = v * (1/r), dot left to right; src/OptiX.cu:62-78,
151-160, 281-318) and the ray is traced by the oracle's any-hit, so the result
is the oracle's definition, bit for bit. A ray never tests the occluder
triangles its own end points lie on.
"""
import numpy as np

F32 = np.float32


def _ray(o, t):
    v = (t - o).astype(F32)
    r = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = (F32(1.0) / r).astype(F32)
        uv = (v * inv[..., None]).astype(F32)
    return uv, r


def _dot(a, n):
    return ((a[..., 0] * n[..., 0] + a[..., 1] * n[..., 1]) + a[..., 2] * n[..., 2]).astype(F32)


def _trace(s, o, uv, tmin, r, ok, skip_a=None, skip_b=None):
    """skip_a, skip_b: occluder triangles of the ray's origin and target (broadcast to ok), or None."""
    hit = np.zeros(ok.shape, bool)
    idx = np.nonzero(ok.ravel())[0]
    if len(idx):
        oo = np.broadcast_to(o, ok.shape + (3,)).reshape(-1, 3)[idx]
        dd = uv.reshape(-1, 3)[idx]
        tm = (r.reshape(-1)[idx] - F32(1e-3)).astype(F32)
        sa = None if skip_a is None else np.broadcast_to(skip_a, ok.shape).reshape(-1)[idx]
        sb = None if skip_b is None else np.broadcast_to(skip_b, ok.shape).reshape(-1)[idx]
        hit.ravel()[idx] = s.anyhit(oo, dd, F32(tmin), tm, skip_a=sa, skip_b=sb)
    return ok & ~hit


def _cge(c):
    with np.errstate(invalid="ignore"):
        return c >= F32(1e-6)


def p2f(s, src, cells_pr, cell_n, use_normal, cell_tri=None):
    """[cell, source]; cell_tri: occluder triangle of each cell (None: not an occluder)."""
    o = src[None, :, :]
    t = cells_pr[:, None, :]
    uv, r = _ray(o, t)
    ok = np.ones(r.shape, bool)
    if use_normal:
        ok &= _cge(_dot(-uv, cell_n[:, None, :]))
    sb = None if cell_tri is None else np.asarray(cell_tri)[:, None]
    return _trace(s, np.broadcast_to(o, uv.shape), uv, 0.0, r, ok, skip_b=sb)


def f2p(s, em_p, em_n, pts, em_tri=None):
    """[point, emitter]: ray from the field point to the fragment."""
    o = pts[:, None, :]
    t = em_p[None, :, :]
    uv, r = _ray(o, t)
    ok = _cge(_dot(-uv, em_n[None, :, :]))
    sb = None if em_tri is None else np.asarray(em_tri)[None, :]
    return _trace(s, np.broadcast_to(o, uv.shape), uv, 0.0, r, ok, skip_b=sb)


def f2f(s, em_p, em_n, rc_p, rc_n, rc_field=False, em_tri=None, rc_tri=None):
    """[receiver, emitter]: ray from the emitter."""
    o = em_p[None, :, :]
    t = rc_p[:, None, :]
    uv, r = _ray(o, t)
    ok = _cge(_dot(uv, em_n[None, :, :]))
    if not rc_field:
        ok &= _cge(_dot(-uv, rc_n[:, None, :]))
    sa = None if em_tri is None else np.asarray(em_tri)[None, :]
    sb = None if rc_tri is None else np.asarray(rc_tri)[:, None]
    return _trace(s, np.broadcast_to(o, uv.shape), uv, 1e-3, r, ok, skip_a=sa, skip_b=sb)


def self_mask(s, p, n, facet, tri_base=0):
    """[receiver, emitter] within one object: ray from the lower facet index."""
    m = np.zeros((len(p), len(p)), bool)
    tri = np.asarray(facet) + tri_base
    fwd = f2f(s, p, n, p, n, em_tri=tri, rc_tri=tri)   # emitter -> receiver
    low_em = facet[None, :] < facet[:, None]
    m[low_em] = fwd[low_em]
    rev = fwd.T                                      # receiver is the origin: transpose of emitter-origin mask
    low_rc = facet[:, None] < facet[None, :]
    m[low_rc] = rev[low_rc]
    return m
