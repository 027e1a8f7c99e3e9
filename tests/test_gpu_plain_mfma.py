"""GPU: the plain FP32 fragment->fragment sums with the complex products on the
FP32 matrix cores (16..64 incidence planes per pass) against the vector-unit
form and the FP64 oracle. Both forms are FP32 end to end; they differ only in
summation order, so they agree to ~1e-6 (bound 1e-5) and both match the oracle to 2e-4."""
import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _run(api, sc, gather_mode):
    runner.load_scene(sc, transfer_mode=0)
    api.set_gather_mode(gather_mode)
    api.profile(True)
    runner.project_scene(sc)
    st = {k: api.kernel_stats(k) for k in ("f2f", "gated_fix")}
    api.profile(False)
    out = runner.read_results(sc)
    api.TearDownCuda()
    return out, st


def _flat(out, p):
    return np.concatenate([out["points"][p], out["targets"][0][p]])


@pytest.mark.parametrize("atten", [0.0, 0.3])
def test_plain_mfma_matches_vector_and_oracle(api, atten):
    sc = scenes.config_trihedral(f=5e3, n_side=4, bounces=3)  # 16 incidences
    sc["atten"] = atten
    _banner(f"plain F2F on matrix cores vs vector units, 16 planes, atten {atten}")
    mf, st_m = _run(api, sc, 0)
    va, st_v = _run(api, sc, 1)
    print("launches mfma", st_m["f2f"]["launches"], "vector", st_v["f2f"]["launches"])
    assert st_m["f2f"]["launches"] == 3 and st_v["f2f"]["launches"] == 3
    e = rel_l2(np.concatenate([mf["points"].ravel(), mf["targets"][0].ravel()]),
               np.concatenate([va["points"].ravel(), va["targets"][0].ravel()]))
    print("matrix cores vs vector units rel-L2", e)
    assert e < 1e-5
    for p in (0, 7, 15):
        ref = oracle.run_scene(sc, plane_source=p)
        b = np.concatenate([ref["points"], ref["targets"][0]])
        print("plane", p, "vs oracle", rel_l2(_flat(mf, p), b))
        assert rel_l2(_flat(mf, p), b) < 2e-4


def test_plain_mfma_gated_rows(api):
    """Incidences from outside the octant leave plates unlit in some planes only."""
    sc = scenes.config_trihedral(f=5e3, n_side=4, bounces=3)
    rng = np.random.default_rng(5)
    src = sc["sources"].copy()
    flip = rng.random(len(src)) < 0.5
    src[flip, 0] *= -1.0  # half the incidences come from x < 0
    sc["sources"] = src.astype(np.float32)
    sc["points"] = sc["sources"].copy()
    _banner("plain F2F on matrix cores, gated rows")
    mf, st_m = _run(api, sc, 0)
    va, _ = _run(api, sc, 1)
    assert st_m["gated_fix"]["launches"] >= 1
    e = rel_l2(np.concatenate([mf["points"].ravel(), mf["targets"][0].ravel()]),
               np.concatenate([va["points"].ravel(), va["targets"][0].ravel()]))
    print("matrix cores vs vector units rel-L2", e)
    assert e < 1e-5
    for p in np.flatnonzero(flip)[:2].tolist() + np.flatnonzero(~flip)[:1].tolist():
        ref = oracle.run_scene(sc, plane_source=p)
        b = np.concatenate([ref["points"], ref["targets"][0]])
        print("plane", p, "vs oracle", rel_l2(_flat(mf, p), b))
        assert rel_l2(_flat(mf, p), b) < 2e-4


def test_plain_mfma_64_planes_and_remainder(api):
    """70 planes: one pass of 64 and one of 6 (the second below the matrix-core threshold is
    still handled by the same pass structure)."""
    sc = scenes.config_trihedral(f=3e3, n_side=8, bounces=2)
    src = np.concatenate([sc["sources"], sc["sources"][:6] * np.float32(1.1)])
    sc["sources"] = src
    sc["points"] = src.copy()
    _banner("plain F2F on matrix cores, 70 planes")
    mf, _ = _run(api, sc, 0)
    va, _ = _run(api, sc, 1)
    e = rel_l2(np.concatenate([mf["points"].ravel(), mf["targets"][0].ravel()]),
               np.concatenate([va["points"].ravel(), va["targets"][0].ravel()]))
    print("matrix cores vs vector units rel-L2", e)
    assert e < 1e-5
    ref = oracle.run_scene(sc, plane_source=69)
    assert rel_l2(_flat(mf, 69), np.concatenate([ref["points"], ref["targets"][0]])) < 2e-4


def test_field_points_on_matrix_cores(api):
    """Fragment->field point for 64 planes on the matrix cores (FP32 within 32-emitter
    splits, FP64 across them) against the compensated vector-unit gather and the oracle."""
    sc = scenes.config_trihedral(f=10e3, n_side=8, bounces=0)
    _banner("F2P on matrix cores, 64 planes, single scattering")
    mf, _ = _run(api, sc, 0)
    va, _ = _run(api, sc, 1)
    e = rel_l2(mf["points"].ravel(), va["points"].ravel())
    print("field points, matrix cores vs vector units rel-L2", e)
    assert e < 1e-5
    for p in (0, 31, 63):
        ref = oracle.run_scene(sc, plane_source=p)
        ep = rel_l2(mf["points"][p], ref["points"])
        print("plane", p, "field points vs oracle", ep)
        assert ep < 1e-4


def test_plain_repeated_scenes_with_side_work(api):
    """Scenes back to back on one model (queued with no host wait between them), the
    visibility words rebuilt every scene, the gated rows' compact sets on the second
    stream beside each bounce's main pass. Every scene equals the first bit for bit."""
    sc = scenes.config_trihedral(f=5e3, n_side=4, bounces=2)
    src = sc["sources"].copy()
    src[::2, 0] *= -1.0  # half the incidences leave a plate unlit: gated rows every bounce
    sc["sources"] = src.astype(np.float32)
    sc["points"] = sc["sources"].copy()
    _banner("plain form: 4 scenes back to back, visibility rebuilt each scene")
    runner.load_scene(sc, transfer_mode=0)
    outs = []
    for _ in range(4):
        api.reset_pressures()
        api.clear_vis_cache()
        runner.project_scene(sc)
        outs.append(runner.read_results(sc))
    api.TearDownCuda()
    for k in range(1, 4):
        assert np.array_equal(outs[k]["points"], outs[0]["points"])
        assert np.array_equal(outs[k]["targets"][0], outs[0]["targets"][0])
    ref = oracle.run_scene(sc, plane_source=0)
    e = rel_l2(_flat(outs[3], 0), np.concatenate([ref["points"], ref["targets"][0]]))
    print("last scene, plane 0 vs oracle", e)
    assert e < 2e-4
