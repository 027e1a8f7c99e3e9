"""GPU: the capped-cylinder hull workload as specified (6 m x 0.5 m hull with
hemispherical caps and a 1 x 0.3 x 0.05 m fin, 8 kHz, 0.5 dB/m, two bounces).

* 720 monostatic angles at 50 m batched as the transfer form's right-hand
  sides, fragments coarsened to lambda/4 with the learned near-field correction
  in the attenuating medium: seeded angles against the FP64 oracle with its
  16 x 16 near-field judge at the same fragments (the oracle's sums carry
  e^{-alpha r} exactly as the reference's thrust::exp(i k r) with Im k = alpha,
  src/OptiX.cu:350, src/Model.cpp:52-53);
* the same at lambda/7.5 in the plain and transfer forms against the oracle's
  point rule (the reference's own discretisation);
* one 512 x 512 field-surface render (FACET_NO_RESP, src/OptiX.cu:192-406):
  seeded cells against the oracle. With the plane 1.5 m or more from every
  fragment no cell sees the r < delta branch, so a field-surface cell is 0.9
  times the fragment->point sum at its point (OptiX.cu:182 vs :340-352), which
  the oracle evaluates there with the same shadow rule.
"""
import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SEEDED_ANGLES = (0, 197, 441, 719)


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _flat(out, p):
    return np.concatenate([out["points"][p], out["targets"][0][p]])


def _run(api, sc, mode, near=None):
    runner.load_scene(sc, transfer_mode=mode)
    if near:
        api.set_near_mode(near[0], near[1])
    runner.project_scene(sc)
    out = runner.read_results(sc)
    api.TearDownCuda()
    return out


def test_hull_lambda4_learned_attenuated_720_angles(api):
    sc = scenes.config_hull(f=8e3, frac=4.0, n_angles=720)
    lam = 1480.0 / 8e3
    assert sc["atten"] > 0
    _banner("hull 8 kHz, lambda/4, 720 angles, 0.5 dB/m, learned near field vs the oracle's judge")
    api.nf_init(1)
    print("train / held-out", api.nf_train(nsamples=262144, steps=20000, batch=4096, lr=3e-3, seed=7), flush=True)
    ml = _run(api, sc, 1, (api.NEAR_LEARNED, 2.0))
    qd = _run(api, sc, 1, (api.NEAR_QUADRATURE, 2.0))
    assert ml["points"].shape == (720, 720)
    for p in SEEDED_ANGLES:
        ref = oracle.run_scene(sc, plane_source=p, near=(16, 2 * lam))
        b = np.concatenate([ref["points"], ref["targets"][0]])
        e_ml, e_qd = rel_l2(_flat(ml, p), b), rel_l2(_flat(qd, p), b)
        print("angle", p, "learned vs judge", e_ml, "quadrature vs judge", e_qd, flush=True)
        assert e_qd < 1e-3 and e_ml < 1e-3
    assert np.all(np.isfinite(ml["points"]))


def test_hull_lambda7p5_720_angles_plain_and_transfer(api):
    sc = scenes.config_hull(f=8e3, n_angles=720)
    _banner("hull 8 kHz, lambda/7.5 (63,764 fragments), 720 angles: plain and transfer vs the oracle")
    pl = _run(api, sc, 0)
    tr = _run(api, sc, 1)
    for p in SEEDED_ANGLES[:2]:
        ref = oracle.run_scene(sc, plane_source=p)
        b = np.concatenate([ref["points"], ref["targets"][0]])
        e_pl, e_tr = rel_l2(_flat(pl, p), b), rel_l2(_flat(tr, p), b)
        print("angle", p, "plain", e_pl, "transfer", e_tr, flush=True)
        assert e_pl < 2e-4 and e_tr < 1e-3


def test_hull_field_surface_512(api):
    sc = scenes.config_hull(f=8e3, n_angles=1, field_cells=512)
    # the plane holds the scattered field only (no direct source term on it)
    sc = dict(sc, batched=False, project_field_objects=False)
    _banner("hull 8 kHz, one incidence, 512 x 512 field surface: seeded cells vs the oracle")
    runner.load_scene(sc)
    runner.project_scene(sc)
    surf = api.object_pressures(2, 0, 0, 0)
    act = api.active_cells(2, 0)
    tgt = api.object_pressures(0, 0, 0, 0)
    api.TearDownCuda()
    # a square of 512 fragment edges a side, split into two triangles: ~512^2 fragments
    assert abs(len(act) / 512 ** 2 - 1) < 0.02, len(act)
    rng = np.random.default_rng(0)
    pick = act[np.sort(rng.choice(len(act), 256, replace=False))]
    # the oracle at the picked cells' fragment points (the field plane's emitter-form points)
    s = oracle.Scene()
    tris, d = sc["fields"][0]
    s.load_geometry(tris, 2, d)
    s.render()
    cells = s.cells(2, 0)
    s.close()
    ref = oracle.run_scene(dict(sc, fields=[], project_field_objects=False, points=cells["pe"][pick]))
    assert rel_l2(tgt, ref["targets"][0]) < 2e-4
    want = 0.9 * ref["points"]
    e = rel_l2(surf[pick], want)
    print("field surface cells vs 0.9 x oracle fragment->point sums", e, "max |p|", np.abs(want).max(), flush=True)
    assert e < 2e-4
    assert np.count_nonzero(surf) > 0.3 * len(act)
