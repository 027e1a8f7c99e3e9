"""GPU: batched uneven scenes in persistent multi-scene launches and one hipGraph.

Every scene of a batch must equal the FP64 oracle run of that scene alone;
graph replays (with eager work between them) must equal the eager run bit for
bit; the autotuner must leave results unchanged.
"""
import numpy as np
import pytest

import oracle
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _small_batch():
    b = scenes.config_batch(n_scenes=10, seed=3, n_points=16)
    for sc in b:  # keep the oracle quick
        sc["f"] = min(sc["f"], 6e3)
        sc["targets"] = [(sc["targets"][0][0], sc["c"] / sc["f"] / 4.0)]
    tri = scenes.config_trihedral(f=4e3, frac=5.0, n_side=1, bounces=2)
    tri["points"] = scenes.generate_field_points(10.0, np.linspace(-60, 60, 16))
    tri["batched"] = False
    plate = scenes.make_rectangle(0.8, 0.8)
    wall = plate.copy()
    wall[:, :, [1, 2]] = wall[:, :, [2, 1]]
    wall[:, :, 1] += 0.4
    wall[:, :, 2] += 0.4
    vee = dict(name="vee", f=5e3, c=1480.0, atten=0.2, targets=[(np.concatenate([plate, wall]), 1480 / 5e3 / 5)],
               fields=[], sources=np.array([[0.5, -3.0, 4.0]], np.float32),
               points=scenes.generate_field_points(8.0, np.linspace(-50, 50, 16)), bounces=2)
    return b + [tri, vee]


def _load(api, batch, near=0):
    api.batch_clear()
    api.batch_set_bounces(2)
    api.batch_set_near(near)
    for sc in batch:
        tris, d = sc["targets"][0]
        api.batch_add_scene(tris, d, sc["f"], sc["sources"], sc["points"], c=sc["c"], atten=sc.get("atten", 0.0))
    api.batch_prepare()


def test_batch_scenes_match_oracle(api):
    batch = _small_batch()
    _banner(f"batch of {len(batch)} uneven scenes vs oracle")
    _load(api, batch)
    api.batch_run(use_graph=False)
    worst = 0.0
    for i, sc in enumerate(batch):
        sc = dict(sc, bounces=2)
        ref = oracle.run_scene(sc)["points"]
        got = api.batch_field_points(i, len(sc["points"]))
        e = rel_l2(got, ref)
        print(sc["name"], "rel-L2", e, "|p| max", np.abs(ref).max())
        worst = max(worst, e)
    api.batch_clear()
    assert worst < 2e-4


def test_batch_graph_replays_match_eager(api):
    batch = _small_batch()
    _banner("batch: hipGraph replays vs eager")
    _load(api, batch)
    api.batch_run(use_graph=False)
    eager = [api.batch_field_points(i, len(sc["points"])) for i, sc in enumerate(batch)]
    for rep in range(3):
        api.batch_run(use_graph=True)
        got = [api.batch_field_points(i, len(sc["points"])) for i, sc in enumerate(batch)]
        for a, b in zip(got, eager):
            assert np.array_equal(a, b)
        api.batch_run(use_graph=False)  # eager work between replays
    api.batch_clear()


def test_batch_autotune_keeps_results(api):
    batch = _small_batch()
    _banner("batch autotune")
    _load(api, batch)
    api.batch_run(use_graph=False)
    before = [api.batch_field_points(i, len(sc["points"])) for i, sc in enumerate(batch)]
    rep = api.batch_autotune()
    print("candidates (ms, visGrid, gatherGrid, emPerWave, minPerSplit):")
    print(rep)
    print("chosen", api.batch_tune())
    assert len(rep) >= 8
    api.batch_run(use_graph=True)
    after = [api.batch_field_points(i, len(sc["points"])) for i, sc in enumerate(batch)]
    for a, b in zip(after, before):
        assert rel_l2(a, b) < 1e-6
    api.batch_clear()


@pytest.mark.parametrize("near", [1, 2])
def test_batch_near_field_vs_oracle_judge(api, near):
    """Near pairs (< 2 lambda) by the 16x16 quadrature (1) or the learned correction (2) in
    the batched bounces: every scene against the oracle with its 16x16 near-field judge;
    graph replays equal the eager run bit for bit."""
    batch = _small_batch()
    _banner(f"batch near field mode {near} vs the oracle's judge")
    if near == 2:
        api.nf_init(1)
        print("train / held-out", api.nf_train(nsamples=262144, steps=20000, batch=4096, lr=3e-3, seed=7))
    _load(api, batch, near)
    api.batch_run(use_graph=False)
    eager = [api.batch_field_points(i, len(sc["points"])) for i, sc in enumerate(batch)]
    worst, moved = 0.0, 0.0
    for i, sc in enumerate(batch):
        sc = dict(sc, bounces=2)
        lam = sc["c"] / sc["f"]
        ref = oracle.run_scene(sc, near=(16, 2 * lam))["points"]
        pt = oracle.run_scene(sc)["points"]
        e = rel_l2(eager[i], ref)
        print(sc["name"], "vs judge", e, "judge vs point rule", rel_l2(ref, pt))
        worst = max(worst, e)
        moved = max(moved, rel_l2(ref, pt))
    assert worst < 1e-3
    assert moved > 1e-3  # the near field matters in this batch (the concave scenes)
    api.batch_run(use_graph=True)
    for i, sc in enumerate(batch):
        assert np.array_equal(api.batch_field_points(i, len(sc["points"])), eager[i])
    api.batch_clear()


def test_batch_256_scenes_learned_lambda4(api):
    """The batched workload as specified: 256 seeded uneven scenes (convex polyhedra of
    8-40 triangles or plates, 0.3-1.5 m, 2-20 kHz, one source, 32 field points each, 2
    bounces) at lambda/4 with the learned near-field correction, replayed as one hipGraph
    after autotuning. A seeded subset of 16 scenes is checked against the FP64 oracle with
    its 16x16 near-field judge (the whole batch is checked by bench.py --config batch)."""
    batch = scenes.config_batch()
    assert len(batch) == 256
    _banner("256-scene batch, learned lambda/4 near field, hipGraph, vs the oracle's judge on 16 scenes")
    api.nf_init(1)
    print("train / held-out", api.nf_train(nsamples=262144, steps=20000, batch=4096, lr=3e-3, seed=7), flush=True)
    _load(api, batch, near=2)
    api.batch_autotune()
    api.batch_run(use_graph=True)
    pick = np.sort(np.random.default_rng(11).choice(len(batch), 16, replace=False))
    got, want = [], []
    for i in pick:
        sc = batch[i]
        lam = sc["c"] / sc["f"]
        ref = oracle.run_scene(sc, near=(16, 2 * lam))["points"]
        g = api.batch_field_points(int(i), len(sc["points"]))
        print(sc["name"], f"{sc['f']:.0f} Hz", "rel-L2 vs judge", rel_l2(g, ref), flush=True)
        assert rel_l2(g, ref) < 5e-3
        got.append(g)
        want.append(ref)
    total = rel_l2(np.concatenate(got), np.concatenate(want))
    print("subset rel-L2", total, flush=True)
    assert total < 1e-3
    api.batch_clear()
