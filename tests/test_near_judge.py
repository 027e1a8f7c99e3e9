"""Pins the oracle's near-field judge (16 x 16 sub-fragment quadrature, CPU only).

The judge replaces the reference's point rule A cos_s cos_i 0.9 G(r)
(src/OptiX.cu:340-354) by the mean of the same kernel over nsub x nsub
sub-points of the emitter's delta x delta cell. Three independent checks:
  * nsub = 1 is the point rule itself: a whole scene run with the judge on every
    pair at nsub = 1 equals the oracle's point-rule run (its other code path,
    pinned by the dense-matrix and closed-form tests of test_oracle.py) to FP32
    geometry rounding;
  * far away, the sub-point phases e^{-ik u.s} of a symmetric n x n grid sum to
    the discrete Dirichlet kernel: quad / point -> D_n(k delta u_x) D_n(k delta u_y)
    with D_n(a) = sin(a/2) / (n sin(a/(2n))) and u_x, u_y the direction's
    components along the emitter facet's axes;
  * near pairs (including one whose receiver plane cuts the emitter cell) agree
    with scipy's adaptive dblquad of the continuous integrand over the cell
    within the midpoint rule's error.
"""
import cmath
import math

import numpy as np
import pytest
from scipy import integrate

import oracle
import scenes
from conftest import rel_l2

C = 1480.0


def _kernel(k, pe, xa, ya, na, nb, pr, ox, oy, h):
    """Continuous integrand at cell offset (ox, oy): cos_s cos_i 0.9 G(r) (FP64)."""
    p = pe + ox * xa + oy * ya
    v = pr - p
    r = math.sqrt(v @ v)
    cs = (v @ na) / r
    ci = 1.0 if nb is None else -(v @ nb) / r
    if not (cs >= 1e-6 and ci >= 1e-6):
        return 0j
    if r < h:
        return 0.9 * cs * ci * (1j * k / (4 * math.pi))
    G = -np.exp(1j * k * r) / (4 * math.pi) * (1j * k / r) + 1.0 / (r * r)
    return 0.9 * cs * ci * G


def test_nsub1_equals_point_rule_scene():
    """nsub = 1 on every pair reproduces the oracle's point-rule run (trihedral, 2 bounces)."""
    sc = scenes.config_trihedral(f=5e3, n_side=1, bounces=2)
    ref = oracle.run_scene(sc, plane_source=0)
    got = oracle.run_scene(sc, plane_source=0, near=(1, 100.0))
    a = np.concatenate([got["points"], got["targets"][0]])
    b = np.concatenate([ref["points"], ref["targets"][0]])
    # FP64 geometry in the judge vs the point rule's FP32 r, cosines and 1/r^2
    assert rel_l2(a, b) < 2e-6, rel_l2(a, b)
    # and the judge at 16 x 16 moves the result by much more than that rounding
    q16 = oracle.run_scene(sc, plane_source=0, near=(16, 2 * C / 5e3))
    assert rel_l2(np.concatenate([q16["points"], q16["targets"][0]]), b) > 1e-3


def _dirichlet(a, n):
    """sum over the n sub-points of e^{-i a s} (s the sub-point offsets in cells), / n; a may be complex."""
    return 1.0 if abs(a) < 1e-12 else cmath.sin(a / 2) / (n * cmath.sin(a / (2 * n)))


@pytest.mark.parametrize("u,frac,att", [((0.7, 0.2, 0.6), 4.0, 0.0), ((0.1, -0.8, 0.55), 4.0, 0.0),
                                        ((-0.5, 0.4, 0.75), 7.5, 0.0), ((0.6, 0.3, 0.5), 4.0, 0.05)])
def test_far_field_dirichlet_closed_form(u, frac, att):
    f = 20e3
    lam = C / f
    kr = float(np.float32(2 * np.pi * f) / np.float32(C))
    d = lam / frac
    # an oblique emitter frame: the axes are not the coordinate axes
    rng = np.random.default_rng(3)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    xa, ya, na = q[:, 0], q[:, 1], np.cross(q[:, 0], q[:, 1])
    uu = np.asarray(u, float)
    uu /= np.linalg.norm(uu)
    dirw = uu[0] * xa + uu[1] * ya + uu[2] * na
    pe = np.array([0.3, -0.2, 0.1])
    R = 3000 * d
    pr = pe + R * dirw
    nb = -dirw  # receiver facing the emitter: cos_i = 1 at every sub-point to O(delta / R)
    vals = {}
    for nsub in (1, 16):
        vals[nsub] = oracle.near_pair(kr, att, nsub, d, d * d, pe, xa, ya, na, nb, pr)
    kd = complex(kr, att) * d  # attenuation: k complex, the same sum
    # the phase of sub-point s is e^{-i k u.s}: Dirichlet factors along each axis. They
    # multiply the e^{ikr} part of G only; the real +1/r^2 term of the reference
    # (OptiX.cu:350, outside the exponential) is smooth across the cell and keeps its
    # point value: 0.9 A cos_s cos_i / R^2 with cos_s = u_z, cos_i = 1.
    T2 = 0.9 * d * d * uu[2] / (R * R)
    D = _dirichlet(kd * uu[0], 16) * _dirichlet(kd * uu[1], 16)
    want = (D * (vals[1] - T2) + T2) / vals[1]
    ratio = vals[16] / vals[1]
    assert abs(ratio - want) < 5e-5, (ratio, want)
    # the check resolves the cell's plane: dropping the phase (ratio 1) or laying the
    # sub-grid along the normal instead of the y axis gives another answer
    assert abs(want - 1.0) > 1e-2
    assert abs(ratio - D) > 5e-5 or abs(T2 / vals[1]) < 1e-4  # the 1/r^2 term is resolved
    if abs(abs(uu[1]) - abs(uu[2])) > 0.2:
        wrong_plane = _dirichlet(kd * uu[0], 16) * _dirichlet(kd * uu[2], 16)
        assert abs(ratio - wrong_plane) > 1e-3


@pytest.mark.parametrize("case", ["oblique", "cut", "grazing"])
def test_near_pairs_match_adaptive_quadrature(case):
    f = 20e3
    lam = C / f
    kr = float(np.float32(2 * np.pi * f) / np.float32(C))
    d = lam / 4
    xa, ya, na = np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), np.array([0, 0, 1.0])
    pe = np.zeros(3)
    if case == "oblique":    # receiver 2.5 cells away, facing back at an angle
        pr = np.array([1.2 * d, -0.7 * d, 2.1 * d])
        nb = -np.array([0.3, -0.2, 0.9]) / np.linalg.norm([0.3, -0.2, 0.9])
        cut = None
    elif case == "cut":      # receiver plane x = 0.2 d cuts the emitter cell: only ox < 0.2 d contributes
        pr = np.array([0.2 * d, 0.3 * d, 1.6 * d])
        nb = np.array([-1.0, 0, 0])
        cut = 0.2 * d
    else:                    # near-grazing direction from a receiver at right angles
        pr = np.array([3.0 * d, 0.4 * d, 0.6 * d])
        nb = np.array([-1.0, 0, 0])
        cut = None
    h = d / 16
    got = oracle.near_pair(kr, 0.0, 16, d, d * d, pe, xa, ya, na, nb, pr)
    xhi = d / 2 if cut is None else cut

    def part(fn):
        val, _ = integrate.dblquad(lambda oy, ox: fn(_kernel(kr, pe, xa, ya, na, nb, pr, ox, oy, h)),
                                   -d / 2, xhi, -d / 2, d / 2, epsabs=1e-10, epsrel=1e-10)
        return val

    want = (part(lambda z: z.real) + 1j * part(lambda z: z.imag))  # mean over the cell times its area
    err = abs(got - want) / abs(want)
    # midpoint rule with h = delta/16 on a smooth integrand (or one with a kink where cos_i -> 0
    # at the cut): relative error O(h^2 f''/f); here below 2e-3
    assert err < 2e-3, (case, got, want, err)
    # a half-cell shift of the sub-grid, a typical indexing slip, is far outside that bound
    if case != "cut":
        shifted, _ = integrate.dblquad(
            lambda oy, ox: _kernel(kr, pe, xa, ya, na, nb, pr, ox + h / 2, oy, h).real, -d / 2, d / 2, -d / 2,
            d / 2, epsabs=1e-10, epsrel=1e-10)
        assert abs(shifted - want.real) > 2e-3 * abs(want)
