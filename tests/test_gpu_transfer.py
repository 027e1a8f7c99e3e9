"""GPU: the FP16 MFMA transfer form of the bounce against the FP64 oracle and
against the plain FP32 form. Bound: rel-L2 <= 1e-3 (the metric's accuracy bar)."""
import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_MIXED = 1e-3


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def test_mixed_trihedral_vs_oracle_gated(api):
    sc = scenes.config_trihedral(f=5e3, n_side=2, bounces=3)
    _banner("mixed form: trihedral 5 kHz, 4 incidences, 3 bounces")
    out = runner.run_scene(sc, transfer_mode=1)
    worst = 0.0
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p)
        e = rel_l2(np.concatenate([out["points"][p], out["targets"][0][p]]),
                   np.concatenate([ref["points"], ref["targets"][0]]))
        print("plane", p, "rel-L2", e)
        worst = max(worst, e)
    assert worst < TOL_MIXED


def test_transfer_split_counts_agree(api):
    """The product's split-K (each 128-row group's non-zero chunks dealt out evenly
    over the splits, the splits' partials reduced after) gives the same pressures
    for any split count: only the FP32 summation order changes. The 12 kHz
    trihedral's facets each skip a third of K, so the splits' K ranges jump the
    skipped chunks and the boundary of the higher facets falls inside a split."""
    sc = scenes.config_trihedral(f=12e3, n_side=2, bounces=3)
    outs = {}
    for n in (0, 1, 3, 7, 64):
        _banner(f"transfer form, 12 kHz trihedral, {n} splits (0: automatic)")
        runner.load_scene(sc, transfer_mode=1)
        api.set_transfer_splits(n)
        runner.project_scene(sc)
        outs[n] = runner.read_results(sc)
        api.TearDownCuda()
    base = np.concatenate([outs[0]["points"].ravel(), outs[0]["targets"][0].ravel()])
    for n, o in outs.items():
        got = np.concatenate([o["points"].ravel(), o["targets"][0].ravel()])
        print("splits", n, "rel-L2 vs automatic", rel_l2(got, base))
        assert rel_l2(got, base) < 1e-5
    ref = oracle.run_scene(sc, plane_source=1)
    got = np.concatenate([outs[7]["points"][1], outs[7]["targets"][0][1]])
    assert rel_l2(got, np.concatenate([ref["points"], ref["targets"][0]])) < TOL_MIXED


def _symmetric_cases():
    att = scenes.config_trihedral(f=12e3, n_side=2, bounces=3)
    att["atten"] = 0.1
    tris = scenes.halve_facets(scenes.halve_facets(scenes.trihedral(1.0)))
    runs = dict(name="trihedral_48_facets_20kHz", f=20e3, c=1480.0, atten=0.0,
                targets=[(tris, 1480 / 20e3 / 7.5)], fields=[], sources=scenes.octant_incidences(2, 15.0),
                points=scenes.octant_incidences(2, 15.0), bounces=3, batched=True)
    return [("trihedral 20 kHz", scenes.config_trihedral(f=20e3, n_side=2, bounces=3)),
            ("trihedral 12 kHz, attenuation 0.1", att),
            ("trihedral of 48 facets of 342 fragments (each padded to 384), 20 kHz", runs),
            ("three stacked plates, the middle one shadowing, 10 kHz", _stacked_plates())]


def _stacked_plates():
    """Three parallel plates (one triangle each): the lower one (z = 0, normal +z)
    and the upper one (z = 1, normal -z) face each other; the middle one
    (z = 0.5, normal +z, 0.4 m legs) faces the upper one and shadows part of
    the lower-upper pairs, whose rays cross its plane."""
    tris = np.array([[[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]],
                     [[0.1, 0.1, 0.5], [0.5, 0.1, 0.5], [0.1, 0.5, 0.5]],
                     [[0.0, 0.0, 1.0], [0.0, 1.0, 1.0], [1.0, 0.0, 1.0]]], np.float32)
    # low elevations from the sides, between the plates: they light the lower
    # plate (and the middle one, from above its plane) past the upper one
    src = np.array([[-15.0, 0.3, 0.6], [0.4, -15.0, 0.3], [15.0, 0.2, 0.8]], np.float32)
    return dict(name="stacked_plates", f=10e3, c=1480.0, atten=0.0, targets=[(tris, 1480 / 10e3 / 7.5)], fields=[],
                sources=src, points=src, bounces=3, batched=True)


@pytest.mark.parametrize("case", range(4))
def test_symmetric_s_build_matches_full(api, case):
    """S is symmetric bit for bit: exchanging receiver and emitter negates the
    difference vector exactly, so r and the two cosine products come out the
    same, and the visibility words are symmetric. The build that evaluates only
    the 64 x 64 blocks on and above the diagonal and stores each one transposed
    as well must then give the same pressures, bit for bit, as evaluating every
    block, with about half the pair evaluations, and so must that build with
    the object's own visibility evaluated inside it (mode 2, scenes of at most 4
    occluder triangles) rather than read from the visibility words. The cases
    cover the attenuating pair kernel, many facets per plate, each padded (K
    chunks holding padding rows and emitters, coplanar facets' chunks skipped;
    48 triangles: no visibility inside the build), and an occluder whose plane
    separates the two ends of many pairs (their triangle tests run inside the
    build)."""
    name, sc = _symmetric_cases()[case]
    outs, evals = {}, {}
    for sym in (2, 1, 0):
        _banner(f"transfer form, {name}, S build mode {sym}")
        runner.load_scene(sc, transfer_mode=1)
        api.set_transfer_symmetric(sym)
        api.profile(True)
        runner.project_scene(sc)
        evals[sym] = api.kernel_stats("build_s")["flops"]
        vis = api.kernel_stats("vis")["pairs"]
        api.profile(False)
        outs[sym] = runner.read_results(sc)
        api.TearDownCuda()
        print("mode", sym, "pair evaluations", evals[sym], "visibility pairs", vis)
    assert 0 < evals[1] < 0.55 * evals[0]
    assert evals[2] == evals[1]
    a, b, c = (np.concatenate([o["points"].ravel()] + [t.ravel() for t in o["targets"]])
               for o in (outs[2], outs[1], outs[0]))
    assert np.isfinite(a).all() and np.abs(a).max() > 0
    assert b.tobytes() == c.tobytes()
    assert a.tobytes() == c.tobytes()


def test_stacked_plates_shadowing_vs_oracle(api):
    """The scene whose S build tests an occluder inside it, against the FP64
    oracle (brute-force triangle tests) on every incidence."""
    sc = _stacked_plates()
    _banner("transfer form, three stacked plates, visibility inside the S build")
    out = runner.run_scene(sc, transfer_mode=1)
    worst = 0.0
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p, bruteforce=True)
        e = rel_l2(np.concatenate([out["points"][p], out["targets"][0][p]]),
                   np.concatenate([ref["points"], ref["targets"][0]]))
        print("plane", p, "rel-L2", e)
        worst = max(worst, e)
    assert worst < TOL_MIXED


def test_self_words_filled_when_needed_after_symmetric_build(api):
    """The symmetric S build reads only the SELF visibility words at or above the
    diagonal blocks (its rows always precede its emitters), so the transposition
    that fills in the rest waits until a caller needs them. Read back through the
    C ABI, or used by the plain form's bounce in the same scene, they must then
    be the full symmetric set: the same words and pressures as a fresh scene."""
    sc = scenes.config_trihedral(f=12e3, n_side=2, bounces=3)

    def project():
        api.reset_pressures()
        runner.project_scene(sc)
        return runner.read_results(sc)

    _banner("SELF words read after a symmetric S build vs a fresh scene")
    runner.load_scene(sc, transfer_mode=1)
    project()
    words_after = api.mask_words(2, ta=0, ia=0)
    runner.load_scene(sc, transfer_mode=0)
    words_fresh = api.mask_words(2, ta=0, ia=0)
    assert words_after.tobytes() == words_fresh.tobytes()

    _banner("plain bounce after a symmetric S build in the same scene vs a fresh scene")
    runner.load_scene(sc, transfer_mode=1)
    project()
    api.set_transfer_mode(0)
    after = project()
    fresh = runner.run_scene(sc, transfer_mode=0)
    a, b = (np.concatenate([o["points"].ravel()] + [t.ravel() for t in o["targets"]]) for o in (after, fresh))
    assert np.abs(a).max() > 0
    assert a.tobytes() == b.tobytes()
    api.TearDownCuda()


def test_mixed_many_small_facets(api):
    """Facets with fewer than 8 fragments force K steps that straddle facets."""
    tris = scenes.trihedral(1.0)
    for _ in range(3):
        tris = scenes.halve_facets(tris)
    sc = dict(name="trihedral_192_facets", f=3e3, c=1480.0, atten=0.1, targets=[(tris, 1480 / 3e3 / 7.5)], fields=[],
              sources=scenes.octant_incidences(2, 15.0), points=scenes.octant_incidences(2, 15.0), bounces=3,
              batched=True)
    _banner("mixed form: 192-facet trihedral, attenuation, 3 bounces")
    out = runner.run_scene(sc, transfer_mode=1)
    plain = runner.run_scene(sc, transfer_mode=0)
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p)
        a = np.concatenate([out["points"][p], out["targets"][0][p]])
        b = np.concatenate([ref["points"], ref["targets"][0]])
        c = np.concatenate([plain["points"][p], plain["targets"][0][p]])
        print("plane", p, "mixed", rel_l2(a, b), "plain", rel_l2(c, b))
        assert rel_l2(a, b) < TOL_MIXED
        assert rel_l2(c, b) < 2e-4


def test_mixed_equals_plain_12khz_4_incidences(api):
    sc = scenes.config_trihedral(f=12e3, n_side=2, bounces=3)
    _banner("mixed vs plain: trihedral 12 kHz, 4 incidences")
    a = runner.run_scene(sc, transfer_mode=1)
    b = runner.run_scene(sc, transfer_mode=0)
    e = rel_l2(np.concatenate([a["points"].ravel(), a["targets"][0].ravel()]),
               np.concatenate([b["points"].ravel(), b["targets"][0].ravel()]))
    print("mixed vs plain rel-L2", e)
    assert e < TOL_MIXED


@pytest.mark.parametrize("f", [5e3, 2e3])
def test_mixed_gated_rows_exact(api, f):
    """Incidences from outside the octant leave whole plates unlit: their rows have
    P_in == 0 in some planes only. At 5 kHz every plate holds >= 128 fragments: the
    rows of S are padded to single-facet groups and each split's sums are split at
    the group's facet boundary; at 2 kHz (51 per plate) the product keeps two
    accumulators per row. Either way the gate is applied in the reduction (no
    separate pass over the gated rows), exact against the oracle."""
    sc = scenes.config_trihedral(f=f, n_side=2, bounces=3)
    sc["sources"] = np.array([[-6.0, 4.0, 13.0], [5.0, -3.0, 12.0], [4.0, 5.0, 13.0], [3.0, 4.0, -0.5]], np.float32)
    sc["points"] = sc["sources"].copy()
    _banner(f"mixed form: trihedral {f / 1e3:g} kHz, incidences leaving plates unlit, 3 bounces")
    runner.load_scene(sc, transfer_mode=1)
    api.profile(True)
    runner.project_scene(sc)
    fix = api.kernel_stats("gated_fix")
    gemm = api.kernel_stats("gemm")
    api.profile(False)
    out = runner.read_results(sc)
    api.TearDownCuda()
    print("gated-row launches", fix["launches"], "products", gemm["launches"])
    assert fix["launches"] == 0 and gemm["launches"] == 3
    plain = runner.run_scene(sc, transfer_mode=0)
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p)
        a = np.concatenate([out["points"][p], out["targets"][0][p]])
        b = np.concatenate([ref["points"], ref["targets"][0]])
        c = np.concatenate([plain["points"][p], plain["targets"][0][p]])
        print("plane", p, "mixed", rel_l2(a, b), "plain", rel_l2(c, b))
        assert rel_l2(a, b) < TOL_MIXED
        assert rel_l2(c, b) < 2e-4


def test_mixed_two_plane_groups_repeated_scenes(api):
    """80 incidences: every bounce runs two plane groups (64 + 16), and four of the
    incidences leave plates unlit. Each bounce's reduction also finds the column
    scales of the next bounce's input, per group. The same model then renders
    scenes of 3, 1 and 2 bounces back to back (queued with no host wait between
    them, the visibility words rebuilt each time), each equal to that scene on a fresh
    model."""
    sc = scenes.config_trihedral(f=5e3, n_side=9, bounces=3)
    unlit = np.array([[-6.0, 4.0, 13.0], [5.0, -3.0, 12.0], [4.0, 5.0, 13.0], [3.0, 4.0, -0.5]], np.float32)
    sc["sources"] = np.concatenate([sc["sources"][:76], unlit])
    sc["points"] = sc["sources"].copy()
    _banner("mixed form: trihedral 5 kHz, 80 incidences (two plane groups), 3 bounces")
    mixed = runner.run_scene(sc, transfer_mode=1)
    plain = runner.run_scene(sc, transfer_mode=0)
    e = rel_l2(np.concatenate([mixed["points"].ravel(), mixed["targets"][0].ravel()]),
               np.concatenate([plain["points"].ravel(), plain["targets"][0].ravel()]))
    print("mixed vs plain rel-L2", e)
    assert e < TOL_MIXED
    for p in (0, 63, 64, 79):
        ref = oracle.run_scene(sc, plane_source=p)
        a = np.concatenate([mixed["points"][p], mixed["targets"][0][p]])
        b = np.concatenate([ref["points"], ref["targets"][0]])
        print("plane", p, "mixed vs oracle", rel_l2(a, b))
        assert rel_l2(a, b) < TOL_MIXED
    fresh = {3: mixed}
    for nb in (1, 2):
        fresh[nb] = runner.run_scene(dict(sc, bounces=nb), transfer_mode=1)
    _banner("mixed form: scenes of 3, 1, 2 bounces back to back on one model")
    runner.load_scene(sc, transfer_mode=1)
    for nb in (3, 1, 2):
        api.reset_pressures()
        api.clear_vis_cache()  # fresh visibility words and transfer matrix each scene
        runner.project_scene(dict(sc, bounces=nb))
        out = runner.read_results(sc)
        for key in ("points", "targets"):
            x = out[key] if key == "points" else out[key][0]
            y = fresh[nb][key] if key == "points" else fresh[nb][key][0]
            d = rel_l2(x.ravel(), y.ravel())
            print(nb, "bounces", key, "vs fresh model rel-L2", d)
            assert d < 1e-6
    api.TearDownCuda()


def test_mixed_dihedral_parallel_facet_runs(api):
    """Two square plates at right angles, each two triangles with identical normals:
    pairs within a plate always fail one cosine cull, so the product skips the K chunks
    of the whole plate (not only of the row's own triangle), with the zero-input gate
    inside the product. One incidence leaves a plate unlit. Exact against the oracle
    to the transfer form's bound, and the plain form to its own."""
    o = [0.0, 0.0, 0.0]
    tris = np.array([[o, [1, 0, 0], [1, 1, 0]], [o, [1, 1, 0], [0, 1, 0]],
                     [o, [0, 1, 0], [0, 1, 1]], [o, [0, 1, 1], [0, 0, 1]]], dtype=np.float32)
    src = np.array([[6.0, 4.0, 8.0], [8.0, 3.0, 3.0], [-4.0, 5.0, 6.0]], np.float32)
    lam = 1480.0 / 5e3
    sc = dict(name="dihedral_4_facets", f=5e3, c=1480.0, atten=0.0, targets=[(tris, lam / 7.5)], fields=[],
              sources=src, points=src.copy(), bounces=2, batched=True)
    _banner("mixed form: dihedral of two-triangle plates, 5 kHz, 3 incidences, 2 bounces")
    out = runner.run_scene(sc, transfer_mode=1)
    plain = runner.run_scene(sc, transfer_mode=0)
    for p in range(len(src)):
        ref = oracle.run_scene(sc, plane_source=p)
        a = np.concatenate([out["points"][p], out["targets"][0][p]])
        b = np.concatenate([ref["points"], ref["targets"][0]])
        c = np.concatenate([plain["points"][p], plain["targets"][0][p]])
        print("plane", p, "mixed", rel_l2(a, b), "plain", rel_l2(c, b))
        assert rel_l2(a, b) < TOL_MIXED
        assert rel_l2(c, b) < 2e-4


def _reduce_cases():
    tris = scenes.trihedral(1.0)
    for _ in range(3):
        tris = scenes.halve_facets(tris)
    small = dict(name="trihedral_192_facets", f=3e3, c=1480.0, atten=0.1, targets=[(tris, 1480 / 3e3 / 7.5)],
                 fields=[], sources=scenes.octant_incidences(2, 15.0), points=scenes.octant_incidences(2, 15.0),
                 bounces=3, batched=True)
    two = scenes.config_trihedral(f=5e3, n_side=9, bounces=3)
    two["sources"] = two["sources"][:80]
    two["points"] = two["sources"].copy()
    return [("trihedral 20 kHz, 4 incidences (padded single-facet groups, gate at kB)",
             scenes.config_trihedral(f=20e3, n_side=2, bounces=3)),
            ("192-facet trihedral, attenuation (two accumulators per row)", small),
            ("trihedral 5 kHz, 80 incidences (two plane groups per bounce)", two)]


@pytest.mark.parametrize("case", range(3))
def test_reduce_in_product_matches_separate_pass(api, case):
    """The split-K reduction done by each row group's last workgroup of the product
    (its counter: the other splits' partial sums fenced before their counts) sums
    the splits in split order, as the reduction pass does: every pressure the same
    bit for bit, in both gate layouts and with two plane groups per bounce (the
    counters reused by the second launch), over three bounces, and twice in a row
    on one model (the counters back at zero)."""
    name, sc = _reduce_cases()[case]
    outs = {}
    for inprod in (1, 0):
        _banner(f"transfer form, {name}: reduction {'inside the product' if inprod else 'as its own pass'}")
        runner.load_scene(sc, transfer_mode=1)
        api.set_reduce_in_product(inprod)
        runner.project_scene(sc)
        first = runner.read_results(sc)
        api.reset_pressures()
        api.clear_vis_cache()
        runner.project_scene(sc)
        outs[inprod] = runner.read_results(sc)
        api.TearDownCuda()
        a, b = (np.concatenate([o["points"].ravel()] + [t.ravel() for t in o["targets"]]) for o in (first, outs[inprod]))
        assert a.tobytes() == b.tobytes()
    a, b = (np.concatenate([o["points"].ravel()] + [t.ravel() for t in o["targets"]]) for o in (outs[1], outs[0]))
    assert np.isfinite(a).all() and np.abs(a).max() > 0
    assert a.tobytes() == b.tobytes()
