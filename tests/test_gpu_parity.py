"""GPU parity: the HIP engine, called through its C-ABI, against the FP64 oracle.

Visibility words must match the oracle bit for bit (same FP32 arithmetic);
complex pressures must match within the rel-L2 bounds stated per test.
"""
import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu

# Plain FP32 form vs the FP64 oracle.
TOL_FP32 = 2e-4


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def test_plate_single_bounce_field_points(api):
    sc = scenes.config_plate()
    _banner("plate 10 kHz: P2F + F2P")
    ref = oracle.run_scene(sc)
    out = runner.run_scene(sc)
    e = rel_l2(out["points"][0], ref["points"])
    et = rel_l2(out["targets"][0][0], ref["targets"][0])
    print("rel-L2 points", e, "surface", et)
    assert e < TOL_FP32 and et < TOL_FP32


def test_sphere_small_shadowed(api):
    sc = scenes.config_sphere(f=2.5e3, subdiv=3, n_points=90, surface_cells=48)
    _banner("sphere 2.5 kHz subdiv 3: P2F + NO_RESP + F2P with shadowing")
    ref = oracle.run_scene(sc)
    out = runner.run_scene(sc)
    e = rel_l2(out["points"][0], ref["points"])
    et = rel_l2(out["targets"][0][0], ref["targets"][0])
    ef = rel_l2(out["fields"][0][0], ref["fields"][0])
    print("rel-L2 points", e, "surface", et, "field", ef)
    assert e < TOL_FP32 and et < TOL_FP32 and ef < TOL_FP32


def test_trihedral_bounces_batched(api):
    sc = scenes.config_trihedral(f=5e3, n_side=2, bounces=3)
    _banner("trihedral 5 kHz, 4 incidences, 3 bounces, batched planes")
    out = runner.run_scene(sc)
    for p in range(len(sc["sources"])):
        ref = oracle.run_scene(sc, plane_source=p)
        e = rel_l2(out["points"][p], ref["points"])
        et = rel_l2(out["targets"][0][p], ref["targets"][0])
        print("plane", p, "rel-L2 points", e, "surface", et)
        assert e < TOL_FP32 and et < TOL_FP32


# ---------------------------------------------------------------------------
# Visibility words: bit-exact against the oracle's definition.

def _oracle_tables(s, t, i):
    c = s.cells(t, i)
    fr = s.frames(t, i)
    n = fr[c["facet"], 0:3].astype(np.float32)
    act = c["area"] > 1e-6
    return c, n, act


def test_masks_sphere_exact(api):
    import maskref
    sc = scenes.config_sphere(f=2.5e3, subdiv=3, n_points=90, surface_cells=40)
    _banner("sphere masks: P2F, F2P, NO_RESP bit-exact")
    runner.load_scene(sc)
    s = oracle.Scene()
    s.set_initial_conditions(sc["c"], sc["f"])
    s.load_geometry(sc["targets"][0][0], 0, sc["targets"][0][1])
    s.load_geometry(sc["fields"][0][0], 2, sc["fields"][0][1])
    s.load_source_points(sc["sources"])
    s.load_field_points(sc["points"])
    s.render()
    assert np.array_equal(api.vis_params(), s.vis_params())
    c, n, act = _oracle_tables(s, 0, 0)
    cf, nf, actf = _oracle_tables(s, 2, 0)
    ns, nq = len(sc["sources"]), len(sc["points"])
    # point -> fragment, every cell of the target
    ref = maskref.p2f(s, sc["sources"], c["pr"], n, True, cell_tri=c["facet"])
    got = api.mask_bits(0, ns, len(c["area"]), tb=0, ib=0)
    print("P2F visible", ref.sum(), "of", ref.size)
    assert np.array_equal(got, ref)
    # fragment -> point
    ref = maskref.f2p(s, c["pe"][act], n[act], sc["points"], em_tri=c["facet"][act])
    got = api.mask_bits(1, int(act.sum()), nq, ta=0, ia=0)
    print("F2P visible", ref.sum(), "of", ref.size)
    assert ref.sum() > 0 and np.array_equal(got, ref)
    # target -> field object
    ref = maskref.f2f(s, c["pe"][act], n[act], cf["pe"][actf], None, rc_field=True, em_tri=c["facet"][act])
    got = api.mask_bits(3, int(act.sum()), int(actf.sum()), ta=0, ia=0, tb=2, ib=0)
    print("NO_RESP visible", ref.sum(), "of", ref.size)
    assert ref.sum() > 0 and np.array_equal(got, ref)


def test_mask_self_exact(api):
    import maskref
    sc = scenes.config_trihedral(f=5e3, n_side=1, bounces=1)
    _banner("trihedral SELF mask bit-exact")
    runner.load_scene(sc)
    s = oracle.Scene()
    s.set_initial_conditions(sc["c"], sc["f"])
    s.load_geometry(sc["targets"][0][0], 0, sc["targets"][0][1])
    s.load_source_points(sc["sources"])
    s.load_field_points(sc["points"])
    s.render()
    c, n, act = _oracle_tables(s, 0, 0)
    ref = maskref.self_mask(s, c["pe"][act], n[act], c["facet"][act])
    na = int(act.sum())
    got = api.mask_bits(2, na, na, ta=0, ia=0)
    print("SELF visible", ref.sum(), "of", ref.size)
    assert ref.sum() > 0 and np.array_equal(got, ref)


# ---------------------------------------------------------------------------
# More of the reference's behaviours.

def test_attenuation(api):
    sc = scenes.config_trihedral(f=5e3, n_side=1, bounces=2)
    sc["atten"] = 0.4
    sc["batched"] = False
    _banner("trihedral with attenuation 0.4 (Im k), 2 bounces")
    ref = oracle.run_scene(sc)
    out = runner.run_scene(sc)
    e = rel_l2(out["points"][0], ref["points"])
    et = rel_l2(out["targets"][0][0], ref["targets"][0])
    print("rel-L2 points", e, "surface", et)
    assert e < TOL_FP32 and et < TOL_FP32


def test_coherent_sources_and_field_object_incidence(api):
    sc = scenes.config_sphere(f=2.5e3, subdiv=3, n_points=40, surface_cells=32)
    sc["sources"] = np.array([[0.0, 0.0, 20.0], [12.0, 3.0, -9.0]], np.float32)
    sc["project_field_objects"] = True
    _banner("two coherent sources, incident field onto the field object")
    ref = oracle.run_scene(sc)
    out = runner.run_scene(sc)
    for k, (a, b) in {"points": (out["points"][0], ref["points"]), "target": (out["targets"][0][0], ref["targets"][0]),
                      "field": (out["fields"][0][0], ref["fields"][0])}.items():
        e = rel_l2(a, b)
        print(k, e)
        assert e < TOL_FP32


def test_batched_planes_equal_separate_runs(api):
    sc = scenes.config_trihedral(f=5e3, n_side=2, bounces=2)
    _banner("batched planes vs separate GPU runs")
    out = runner.run_scene(sc)
    for p in range(len(sc["sources"])):
        one = dict(sc, sources=sc["sources"][p:p + 1], batched=False)
        o1 = runner.run_scene(one)
        assert rel_l2(out["points"][p], o1["points"][0]) < 1e-6
        assert rel_l2(out["targets"][0][p], o1["targets"][0][0]) < 1e-6


def test_two_targets_bounce_per_object_and_occlude_each_other(api):
    plate = scenes.make_rectangle(1.0, 1.0)
    plate2 = plate.copy()
    plate2[:, :, 2] += 0.6
    plate2 = plate2[:, ::-1, :].copy()  # faces -z
    small = scenes.trihedral(0.5)
    small[:, :, 0] += 0.1
    sc = dict(name="two_targets", f=4e3, c=1480.0, atten=0.0,
              targets=[(np.concatenate([plate, plate2]), 1480 / 4e3 / 6), (small, 1480 / 4e3 / 6)], fields=[],
              sources=np.array([[2.0, 0.5, 8.0]], np.float32),
              points=scenes.generate_field_points(12.0, np.linspace(-60, 60, 25)), bounces=2)
    _banner("two target objects, 2 bounces")
    ref = oracle.run_scene(sc)
    out = runner.run_scene(sc)
    for i in range(2):
        e = rel_l2(out["targets"][i][0], ref["targets"][i])
        print("target", i, e)
        assert e < TOL_FP32
    e = rel_l2(out["points"][0], ref["points"])
    print("points", e)
    assert e < TOL_FP32


def test_resp_between_two_targets(api):
    plate = scenes.make_rectangle(1.0, 1.0)
    top = plate.copy()
    top[:, :, 2] += 0.8
    top = top[:, ::-1, :].copy()
    sc = dict(f=4e3, c=1480.0, targets=[(plate, 1480 / 4e3 / 6), (top, 1480 / 4e3 / 6)],
              sources=np.array([[0.3, 0.2, 0.4]], np.float32), points=np.zeros((1, 3), np.float32))
    _banner("FACET_RESP between two targets")
    runner.load_scene(dict(sc, fields=[], bounces=0))
    api.project_source_points_to_objects(False)
    api.prepare_resp(0)
    api.prepare_resp(1)
    api.project_resp(0, 1)
    got = [api.object_pressures(0, i, 2) for i in range(2)]
    s = oracle.Scene()
    s.set_initial_conditions(sc["c"], sc["f"])
    for tris, d in sc["targets"]:
        s.load_geometry(tris, 0, d)
    s.load_source_points(sc["sources"])
    s.load_field_points(sc["points"])
    s.render()
    s.project_src(False)
    s.prepare_resp(0)
    s.prepare_resp(1)
    s.project_resp(0, 1)
    for i in range(2):
        e = rel_l2(got[i], s.pressure(0, i, 2))
        print("P_out target", i, e)
        assert e < TOL_FP32


def _plate_and_field_plane():
    """A 1 m plate (two triangles) and a vertical field plane through it, reaching
    below it; three incidences: above the plate (the plate shadows the field
    cells below it), below it (every plate cell culled by its cosine) and
    grazing."""
    lam = 1480.0 / 6e3
    src = np.array([[1.5, 0.3, 4.0], [-2.0, 0.5, -3.0], [6.0, -0.2, 0.05]], np.float32)
    return dict(name="plate_and_field_plane", f=6e3, c=1480.0, atten=0.0,
                targets=[(scenes.make_rectangle(1.0, 1.0), lam / 6)],
                fields=[(scenes.make_rectangle(1.6, 1.2, xy_plane=False) + np.float32([0.1, 0.0, -0.2]), lam / 6)],
                sources=src, points=scenes.generate_field_points(10.0, np.linspace(-80, 80, 17)), bounces=1,
                project_field_objects=True, batched=True)


def _inline_cases():
    from test_gpu_transfer import _stacked_plates
    return [("plate and field plane, 6 kHz", _plate_and_field_plane()),
            ("trihedral 12 kHz", scenes.config_trihedral(f=12e3, n_side=2, bounces=2)),
            ("three stacked plates, the middle one shadowing, 10 kHz", _stacked_plates())]


@pytest.mark.parametrize("case", range(3))
def test_p2f_inline_visibility_matches_words(api, case):
    """With one incidence plane per source and at most 4 occluder triangles the
    point->fragment sum evaluates each (source, fragment) pair's visibility itself
    (the receiver's cosine cull, the plane-side candidates, the FP32 triangle
    test of the sources->fragments words, which test_masks_sphere_exact pins bit
    for bit against the oracle's): every pressure must be the same, bit for bit,
    as with the words, with no sources->fragments words traced. The plate case
    also checks the field object's incidence (no cosine cull, a shadow) against
    the FP64 oracle."""
    name, sc = _inline_cases()[case]
    outs, vis = {}, {}
    for inl in (1, 0):
        _banner(f"{name}: point->fragment visibility {'inside the sum' if inl else 'from the words'}")
        runner.load_scene(sc)
        api.set_inline_visibility(inl)
        api.profile(True)
        api.project_source_points_to_objects(sc.get("project_field_objects", False))
        vis[inl] = api.kernel_stats("vis")["pairs"]
        api.profile(False)
        if sc.get("bounces", 0):
            api.project_target_to_target_objects(sc["bounces"])
        if sc.get("fields"):
            api.project_target_to_field_objects()
        api.project_target_to_field_points()
        outs[inl] = runner.read_results(sc)
        api.TearDownCuda()
        print("inline", inl, "visibility pairs of the source step", vis[inl])
    assert vis[1] == 0 and vis[0] > 0
    a, b = (np.concatenate([o["points"].ravel()] + [t.ravel() for t in o["targets"]] +
                           [f.ravel() for f in o["fields"]]) for o in (outs[1], outs[0]))
    assert np.isfinite(a).all() and np.abs(a).max() > 0
    assert a.tobytes() == b.tobytes()
    if case == 0:
        for p in range(len(sc["sources"])):
            ref = oracle.run_scene(sc, plane_source=p)
            for k, (x, y) in {"field": (outs[1]["fields"][0][p], ref["fields"][0]),
                              "target": (outs[1]["targets"][0][p], ref["targets"][0])}.items():
                e = rel_l2(x, y) if np.abs(y).max() > 0 else float(np.abs(x).max())
                print("plane", p, k, e)
                assert e < TOL_FP32
