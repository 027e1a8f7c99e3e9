"""GPU edge cases: tiny scenes (fewer fragments than one wave), ragged
incidence counts around the 16- and 64-plane boundaries of the matrix-core
and transfer passes, and empty receiver sets. Each against the FP64 oracle
(or, for shapes the oracle finishes only slowly, against the plain form that
the oracle pins)."""
import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _corner(leg):
    """Two perpendicular right triangles sharing an edge (normals facing each other's side)."""
    o = [0.0, 0.0, 0.0]
    return np.array([[o, [leg, 0, 0], [0, leg, 0]], [o, [0, leg, 0], [0, 0, leg]]], dtype=np.float32)


def _scene(tris, f, n_src, bounces=2, seed=1):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.2, 1.0, (n_src, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    src = (8.0 * d).astype(np.float32)
    return dict(name="edge", f=f, c=1480.0, atten=0.0, targets=[(tris, 1480.0 / f / 7.5)], fields=[],
                sources=src, points=src.copy(), bounces=bounces, project_field_objects=False, batched=True)


def _run(api, sc, mode, gather_mode=0):
    runner.load_scene(sc, transfer_mode=mode)
    api.set_gather_mode(gather_mode)
    runner.project_scene(sc)
    out = runner.read_results(sc)
    api.TearDownCuda()
    return out


def _flat(out, p):
    return np.concatenate([out["points"][p], out["targets"][0][p]])


@pytest.mark.parametrize("n_src", [1, 15, 17, 33])
def test_tiny_corner_all_forms(api, n_src):
    sc = _scene(_corner(0.25), 6e3, n_src)  # ~50 fragments per facet
    _banner(f"tiny corner, {n_src} incidences: plain (matrix cores, vector units) and transfer vs oracle")
    outs = [_run(api, sc, 0, 0), _run(api, sc, 0, 1), _run(api, sc, 1)]
    for p in sorted({0, n_src // 2, n_src - 1}):
        ref = oracle.run_scene(sc, plane_source=p)
        b = np.concatenate([ref["points"], ref["targets"][0]])
        e = [rel_l2(_flat(o, p), b) for o in outs]
        print("plane", p, "mfma", e[0], "vector", e[1], "transfer", e[2])
        assert e[0] < 2e-4 and e[1] < 2e-4 and e[2] < 1e-3


def test_transfer_two_plane_groups(api):
    """65 incidences: the transfer form's second group of one plane, against the plain form."""
    sc = scenes.config_trihedral(f=4e3, n_side=8, bounces=2)
    sc["sources"] = np.concatenate([sc["sources"], sc["sources"][:1] * np.float32(1.2)])
    sc["points"] = sc["sources"].copy()
    _banner("trihedral 4 kHz, 65 incidences: transfer vs plain, and plane 64 vs the oracle")
    plain = _run(api, sc, 0)
    mixed = _run(api, sc, 1)
    e = rel_l2(np.concatenate([mixed["points"].ravel(), mixed["targets"][0].ravel()]),
               np.concatenate([plain["points"].ravel(), plain["targets"][0].ravel()]))
    ref = oracle.run_scene(sc, plane_source=64)
    e64 = rel_l2(_flat(mixed, 64), np.concatenate([ref["points"], ref["targets"][0]]))
    print("transfer vs plain", e, "plane 64 vs oracle", e64)
    assert e < 1e-3 and e64 < 1e-3


def test_no_field_points_and_no_bounces(api):
    sc = _scene(_corner(0.25), 6e3, 4, bounces=0)
    sc["points"] = np.zeros((0, 3), np.float32)
    _banner("no field points, no bounces: surface pressures only")
    for mode in (0, 1):
        out = _run(api, sc, mode)
        assert out["points"].shape[1] == 0
        ref = oracle.run_scene(sc, plane_source=2)
        e = rel_l2(out["targets"][0][2], ref["targets"][0])
        print("mode", mode, "surface vs oracle", e)
        assert e < 2e-4
