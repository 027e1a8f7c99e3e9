"""Pins the FP64 oracle (CPU only).

Each check is independent of the oracle's loop structure:
  * dense linear algebra: the bounce recursion P = P_inc + sum T^n P_inc with the
    reference's zero-input gating (OptiX.cu:241-245) rebuilt as matrices in numpy;
  * closed forms: far-field plate TES 20 log10(A / 2 lambda) at broadside and the
    sinc pattern of run_scripts/rect_mono.py:11-30;
  * invariants: linearity in the source amplitude, attenuation as exp(-alpha r);
  * visibility: BVH any-hit == brute force over all triangles, bit for bit, and
    analytic occlusion by a sphere.
"""
import math

import numpy as np
import pytest

import oracle
import scenes
from conftest import rel_l2


def _scene(sc, src=None, att=0.0, field_pts=None, bruteforce=False):
    s = oracle.Scene()
    s.set_bruteforce(bruteforce)
    s.set_initial_conditions(sc["c"], sc["f"], att)
    for tris, d in sc["targets"]:
        s.load_geometry(tris, 0, d)
    s.load_source_points(sc["sources"] if src is None else src)
    s.load_field_points(sc["points"] if field_pts is None else field_pts)
    s.render()
    return s


def _dense_transfer(s, delta, att=0.0):
    """Symmetric transfer S (active x active) of one object, built pair by pair in numpy."""
    c = s.cells(0, 0)
    fr = s.frames(0, 0)
    act = np.where(c["area"] > 1e-6)[0]
    A, P, F = c["area"][act].astype(np.float64), c["pe"][act], c["facet"][act]
    N = fr[F, 0:3]
    n = len(act)
    k = complex(s.kf(), att)
    S = np.zeros((n, n), complex)
    ia, ib = np.where(F[:, None] < F[None, :])
    v = (P[ib] - P[ia]).astype(np.float32)
    r = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(np.float32)
    uv = v * (np.float32(1.0) / r)[:, None]
    cs = (uv[:, 0] * N[ia, 0] + uv[:, 1] * N[ia, 1]) + uv[:, 2] * N[ia, 2]
    ci = (-uv[:, 0] * N[ib, 0] + -uv[:, 1] * N[ib, 1]) + -uv[:, 2] * N[ib, 2]
    ok = (cs >= np.float32(1e-6)) & (ci >= np.float32(1e-6))
    hit = s.anyhit(P[ia], uv, np.float32(1e-3), r - np.float32(1e-3), brute=True)
    ok &= ~hit
    rr = r.astype(np.float64)
    G = -np.exp(1j * k * rr) / (4 * np.pi) * (1j * k / rr) + (np.float32(1.0) / (r * r)).astype(np.float64)
    G = np.where(r < np.float32(delta), 1j * k / (4 * np.pi), G)
    val = np.where(ok, 0.9 * G * (cs * ci).astype(np.float64), 0)
    S[ia, ib] = val
    S[ib, ia] = val
    return act, A, F, S


@pytest.mark.parametrize("att", [0.0, 0.35])
def test_bounces_match_dense_matrix_recursion(att):
    sc = scenes.config_trihedral(f=5e3, n_side=2, bounces=3)
    delta = sc["targets"][0][1]
    s = _scene(sc, src=sc["sources"][1:2], att=att)
    s.project_src(False)
    P0 = s.pressure(0, 0)
    act, A, F, S = _dense_transfer(s, delta, att)
    assert np.allclose(S, S.T)
    lower = F[None, :] < F[:, None]  # emitter facet below the receiver's: ungated
    P = P0[act].copy()
    Pout = P.copy()
    for _ in range(3):
        Pin = Pout
        x = A * Pin
        Pout = np.where(lower, S, 0) @ x + (Pin != 0) * (np.where(~lower, S, 0) @ x)
        P = P + Pout
    s.project_target_to_target(3)
    got = s.pressure(0, 0)[act]
    assert (Pin == 0).any(), "the case must exercise the zero-input gate"
    assert rel_l2(got, P) < 1e-7


def _monostatic_tes(f, a, b, R, angles_deg):
    sc = dict(c=1480.0, f=f, targets=[(scenes.make_rectangle(a, b), 1480.0 / f / 7.5)])
    out = []
    for th in angles_deg:
        p = scenes.generate_field_points(R, [th])
        s = _scene(sc, src=p, field_pts=p)
        s.project_src(False)
        s.project_target_to_field_points()
        out.append(20 * np.log10(abs(s.field_points()[0])) + 40 * np.log10(R))
        s.close()
    return np.array(out)


def test_plate_broadside_tes_closed_form():
    f, a, b, R = 10e3, 1.0, 1.0, 200.0
    lam = 1480.0 / f
    tes = _monostatic_tes(f, a, b, R, [0.0])[0]
    assert abs(tes - 20 * np.log10(a * b / (2 * lam))) < 0.05


def test_plate_sinc_pattern():
    """rect_mono.py:26-28 pattern over the main lobe and first sidelobe."""
    f, a, b, R = 10e3, 1.0, 1.0, 200.0
    lam = 1480.0 / f
    ang = np.array([2.0, 4.0, 6.0, 11.0, 13.0])
    tes = _monostatic_tes(f, a, b, R, ang)
    th = np.radians(ang)
    beta = (2 * np.pi / lam) * a * np.sin(th)
    ana = 10 * np.log10(((a * b) / (2 * lam)) ** 2 * (np.sin(beta) / beta) ** 2 * (0.5 + 0.5 * np.cos(2 * th)))
    assert np.max(np.abs(tes - ana)) < 0.3, (tes, ana)


def test_linearity_in_source_amplitude():
    sc = scenes.config_plate(n_points=19)
    s1 = _scene(sc)
    s1.project_src(False)
    s1.project_target_to_field_points()
    s2 = _scene(sc)
    amp = 2.0 - 3.0j
    s2.set_source_amplitude(0, amp)
    s2.project_src(False)
    s2.project_target_to_field_points()
    assert rel_l2(s2.field_points(), amp * s1.field_points()) < 1e-12


def test_attenuation_factor_point_to_fragment():
    sc = scenes.config_plate(n_points=3)
    alpha = 0.5
    s0 = _scene(sc)
    s0.project_src(False)
    s1 = _scene(sc, att=alpha)
    s1.project_src(False)
    c = s0.cells(0, 0)
    src = sc["sources"][0].astype(np.float32)
    v = (c["pr"] - src).astype(np.float32)
    r = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(np.float64)
    p0, p1 = s0.pressure(0, 0), s1.pressure(0, 0)
    nz = np.abs(p0) > 0
    assert np.allclose(p1[nz] / p0[nz], np.exp(-alpha * r[nz]), rtol=1e-9)


def test_point_to_fragment_single_cell_magnitude_and_phase():
    """|P| = cos(theta)/r and arg P = k r (the reference's e^{+ikr} convention)."""
    sc = dict(c=1480.0, f=10e3, targets=[(np.array([[[0, 0, 0], [0.01, 0, 0], [0, 0.01, 0]]], np.float32), 0.02)],
              sources=np.array([[0.0, 0.0, 3.0]], np.float32), points=np.zeros((1, 3), np.float32))
    s = _scene(sc)
    s.project_src(False)
    c = s.cells(0, 0)
    p = s.pressure(0, 0)
    j = int(np.argmax(np.abs(p)))
    v = (c["pr"][j] - sc["sources"][0]).astype(np.float64)
    r = np.linalg.norm(v)
    cosi = -v[2] / r
    assert abs(abs(p[j]) - cosi / r) < 1e-6 * abs(p[j])
    ph = np.angle(p[j] * np.exp(-1j * s.kf() * r))
    assert abs(ph) < 1e-4


def test_bvh_equals_bruteforce():
    tris, _ = scenes.icosphere(3)
    sc = dict(c=1480.0, f=1e3, targets=[(tris, 0.2)], sources=np.zeros((1, 3), np.float32),
              points=np.zeros((1, 3), np.float32))
    s = _scene(sc)
    rng = np.random.default_rng(7)
    n = 200000
    o = rng.uniform(-2.5, 2.5, (n, 3)).astype(np.float32)
    t = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    v = (t - o).astype(np.float32)
    r = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(np.float32)
    d = v * (np.float32(1) / r)[:, None]
    tmin = np.where(rng.random(n) < 0.5, np.float32(0), np.float32(1e-3)).astype(np.float32)
    tmax = (r - np.float32(1e-3)).astype(np.float32)
    hb = s.anyhit(o, d, tmin, tmax, brute=False)
    hf = s.anyhit(o, d, tmin, tmax, brute=True)
    assert hf.any() and (~hf).any()
    assert np.array_equal(hb, hf)


def test_sphere_analytic_occlusion():
    tris, _ = scenes.icosphere(3)
    sc = dict(c=1480.0, f=1e3, targets=[(tris, 0.2)], sources=np.zeros((1, 3), np.float32),
              points=np.zeros((1, 3), np.float32))
    s = _scene(sc)
    rng = np.random.default_rng(3)
    u = rng.normal(size=(500, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    o = (3.0 * u).astype(np.float32)
    # through the centre to the far side: blocked
    d = (-u).astype(np.float32)
    assert s.anyhit(o, d, 0.0, np.float32(5.9)).all()
    # stopping short of the sphere: clear
    assert not s.anyhit(o, d, 0.0, np.float32(1.9)).any()
    # passing outside the sphere (offset 1.2 m from the centre): clear
    w = np.cross(u, rng.normal(size=(500, 3)))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    o2 = (3.0 * u + 1.2 * w).astype(np.float32)
    assert not s.anyhit(o2, d, 0.0, np.float32(6.0)).any()
