"""GPU: the learned near-field correction.

* the GPU's quadrature labels equal an FP64 numpy quadrature rebuilt from the features;
* the GPU's backward pass equals FP64 back-propagation (gradient check);
* training on the GPU drives the held-out error down;
* in the transfer form, the quadrature mode matches the oracle's near-field
  judge, and the learned mode is closer to that judge than the point rule.
"""
import numpy as np
import pytest

import nf_ref
import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def test_labels_match_fp64_quadrature(api):
    _banner("near-field labels vs FP64 quadrature")
    X, Y = api.nf_samples(11, 256)
    ref = np.array([nf_ref.label(x) for x in X.astype(np.float64)])
    scale = np.abs(ref).max()
    err = np.abs(Y - ref).max() / scale
    print("max label error / max label", err, "label scale", scale)
    assert err < 2e-3


def test_gradient_check_against_fp64(api):
    _banner("MLP gradient check")
    rng = np.random.default_rng(5)
    p = (rng.normal(size=nf_ref.NPAR) * 0.2).astype(np.float32)
    X, Y = api.nf_samples(3, 512)
    g = api.nf_gradient(p, X, Y)
    gr = nf_ref.gradient(p, X, Y)
    e = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print("gradient rel error", e)
    assert e < 1e-4
    out = api.nf_forward(X) if False else None  # forward of the trained net is covered below
    # finite differences on a few parameters pin the FP64 restatement itself
    def loss(q):
        return float(((nf_ref.forward(q, X)[2] - Y) ** 2).sum() / len(X))
    for idx in [3, 600, 4700, 4929]:
        e1 = np.zeros(nf_ref.NPAR)
        e1[idx] = 1e-6
        fd = (loss(p + e1) - loss(p - e1)) / 2e-6
        assert abs(fd - gr[idx]) < 1e-5 * max(1.0, abs(gr[idx]))


def test_training_reduces_heldout_error(api):
    _banner("MLP training on the GPU")
    api.nf_init(1)
    X, Y = api.nf_samples(99, 4096)
    e0 = np.linalg.norm(api.nf_forward(X) - Y) / np.linalg.norm(Y)
    tr, va = api.nf_train(nsamples=131072, steps=8000, batch=4096, lr=3e-3, seed=7)
    e1 = np.linalg.norm(api.nf_forward(X) - Y) / np.linalg.norm(Y)
    print("untrained", e0, "trained: train", tr, "held-out", va, "fresh", e1)
    # the targets are the base rule's remainder scaled to order one, so an
    # untrained network sits near 1; 8000 steps explain most of the variance
    assert va < 0.8 and e1 < 0.8 and e1 < 0.7 * e0
    assert abs(va - e1) < 0.1  # held-out and fresh samples agree: no over-fitting


def _trihedral_near_case():
    sc = scenes.config_trihedral(f=5e3, frac=4.0, n_side=2, bounces=2)
    lam = 1480.0 / 5e3
    return sc, lam


def _mixed_run(api, sc, near_mode):
    runner.load_scene(sc, transfer_mode=1)
    api.set_near_mode(near_mode, 2.0)
    runner.project_scene(sc)
    out = runner.read_results(sc)
    api.TearDownCuda()
    return out


def _flat(out, p):
    return np.concatenate([out["points"][p], out["targets"][0][p]])


def test_quadrature_mode_matches_oracle_judge(api):
    sc, lam = _trihedral_near_case()
    _banner("transfer form with 16x16 near-field quadrature vs oracle judge")
    out = _mixed_run(api, sc, api.NEAR_QUADRATURE)
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p, near=(16, 2 * lam))
        e = rel_l2(_flat(out, p), np.concatenate([ref["points"], ref["targets"][0]]))
        print("plane", p, "quadrature vs judge", e)
        assert e < 1e-3


def test_learned_mode_gain_over_point_rule(api):
    """Against the judge: the learned mode (2 x 2 Gauss base rule + trained
    network) beats the point rule, and the trained network beats its own base
    rule alone (the same mode with every weight zero, so the network adds 0)."""
    sc, lam = _trihedral_near_case()
    _banner("learned near field vs point rule and vs its base rule, against the judge")
    api.nf_init(1)
    print("train / held-out error", api.nf_train(nsamples=262144, steps=30000, batch=4096, lr=3e-3, seed=7))
    trained = api.nf_params()
    assert np.all(np.isfinite(trained))
    pt = _mixed_run(api, sc, api.NEAR_POINT)
    ml = _mixed_run(api, sc, api.NEAR_LEARNED)
    api.nf_set_params(np.zeros_like(trained))
    base = _mixed_run(api, sc, api.NEAR_LEARNED)
    api.nf_set_params(trained)
    gains, net_gains = [], []
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p, near=(16, 2 * lam))
        b = np.concatenate([ref["points"], ref["targets"][0]])
        e_pt, e_ml, e_b = rel_l2(_flat(pt, p), b), rel_l2(_flat(ml, p), b), rel_l2(_flat(base, p), b)
        print("plane", p, "point rule", e_pt, "base rule", e_b, "learned", e_ml)
        gains.append(e_pt / e_ml)
        net_gains.append(e_b / e_ml)
    assert min(gains) > 1.5
    assert min(net_gains) > 1.2
