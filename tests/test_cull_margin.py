"""Pins the margin rule that decides cosine culls without the ray set-up (CPU only).

The symmetric S build evaluates the visibility of an object's own fragment
pairs itself. It decides each cosine cull (the pair is kept when cos >= 1e-6
with cos formed from the unit direction, OptiX.cu:326-333) from x = v.n and
r2 = v.v alone where rounding cannot change the answer (vis_fp32.h
cull_margin): x <= 0 or x^2 < 1e-14 r2 -> culled, x^2 > 9e-12 r2 -> kept,
anything else (and r2 outside [2^-60, 2^60]) -> the exact set-up. These tests
check the rule against the oracle's own FP32 cull (orc_cosine_kept: ray_dir and
dotp as the sums form them) on random pairs and on pairs whose cosine grazes
the threshold, with x and r2 formed both with and without fused
multiply-adds: every decided pair agrees with the oracle, only pairs near the
threshold stay undecided, and a rule without the margin (deciding at exactly
1e-6) would disagree (so the margin is needed, and the grazing pairs find that).
"""
import numpy as np

import oracle

F32 = np.float32


def _fma(a, b, c):
    """FP32 fused multiply-add: the product is exact in FP64, one rounding to FP32
    after the add (the FP64 add's own rounding is far below FP32's)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def _dot_plain(a, b):
    return ((a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]).astype(F32) + a[:, 2] * b[:, 2]).astype(F32)


def _dot_fma(a, b):
    return _fma(a[:, 2], b[:, 2], _fma(a[:, 1], b[:, 1], (a[:, 0] * b[:, 0]).astype(F32)))


def _rule(x, r2, hi=F32(9e-12), lo=F32(1e-14)):
    """cull_margin: 1 kept, 0 culled, 2 undecided."""
    out = np.full(len(x), 2, np.int8)
    inr = (r2 >= F32(2.0 ** -60)) & (r2 <= F32(2.0 ** 60))
    x2 = (x * x).astype(F32)
    kept = inr & (x > 0) & (x2 > (hi * r2).astype(F32))
    culled = inr & (~(x > 0) | (x2 < (lo * r2).astype(F32)))
    out[culled] = 0
    out[kept & ~culled] = 1
    return out


def _unit(rng, k):
    n = rng.normal(size=(k, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return n.astype(F32)


def _cases():
    rng = np.random.default_rng(5)
    k = 400000
    # random pairs at distances 1e-4 .. 1e3
    n1 = _unit(rng, k)
    v1 = (_unit(rng, k).astype(np.float64) * 10.0 ** rng.uniform(-4, 3, (k, 1))).astype(F32)
    # grazing pairs: cosine +-1e-9 .. 3e-5 (dense around 1e-6) against the normal
    n2 = _unit(rng, k)
    t = rng.normal(size=(k, 3))
    t -= (t * n2).sum(1, keepdims=True) * n2
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    c = np.concatenate([10.0 ** rng.uniform(-9, -4.5, k // 2), 1e-6 * (1 + rng.uniform(-0.9, 3.0, k - k // 2))])
    c *= rng.choice([-1.0, 1.0], size=k, p=[0.2, 0.8])
    s = 10.0 ** rng.uniform(-3, 2, (k, 1))
    v2 = (s * (c[:, None] * n2 + np.sqrt(1 - c * c)[:, None] * t)).astype(F32)
    return np.concatenate([v1, v2]), np.concatenate([n1, n2]), k


def test_decided_culls_match_the_oracle():
    v, n, k = _cases()
    kept = oracle.cosine_kept(v, n)
    for dot in (_dot_plain, _dot_fma):
        # both ends of a pair: the emitter's cull uses +v, the receiver's -v
        for sgn in (F32(1), F32(-1)):
            vv = (sgn * v).astype(F32)
            want = kept if sgn > 0 else oracle.cosine_kept(vv, n)
            d = _rule(dot(vv, n), dot(vv, vv))
            assert not np.any((d == 1) & ~want), "a pair the rule keeps is culled by the oracle"
            assert not np.any((d == 0) & want), "a pair the rule culls is kept by the oracle"
            # random pairs are decided; only grazing ones stay undecided, and they do get tested
            assert np.mean(d[:k] == 2) < 1e-4
            assert 0.05 < np.mean(d[k:] == 2) < 0.6
            assert np.any((d[k:] != 2) & want[k:]) and np.any((d[k:] != 2) & ~want[k:])


def test_the_margin_is_needed():
    """Deciding at exactly 1e-6 (no margin) from x and r2 disagrees with the oracle
    on some grazing pairs: the exact set-up's rounding moves the cosine by more
    than the gap."""
    v, n, k = _cases()
    kept = oracle.cosine_kept(v, n)
    d = _rule(_dot_fma(v, n), _dot_fma(v, v), hi=F32(1e-12), lo=F32(1e-12))
    wrong = int(np.sum((d == 1) & ~kept) + np.sum((d == 0) & kept))
    assert wrong > 0
