"""GPU: the capped-cylinder hull with a fin (a 924- or 3756-triangle mesh, so
the visibility walks a real LBVH), attenuated medium, two bounces, many
monostatic incidences batched as planes.

* the SELF visibility words equal the oracle's bit for bit;
* plain FP32 and FP16 transfer forms against the FP64 oracle at 2 kHz;
* at the full 8 kHz mesh (63,764 fragments) the transfer form against the
  plain form (the plain form being the one pinned to the oracle above).
"""
import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SMALL = dict(n_around=24, n_along=12, n_cap=4)


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _flat(out, p):
    return np.concatenate([out["points"][p], out["targets"][0][p]])


def _run(api, sc, mode):
    runner.load_scene(sc, transfer_mode=mode)
    runner.project_scene(sc)
    out = runner.read_results(sc)
    api.TearDownCuda()
    return out


def test_hull_self_mask_exact(api):
    import maskref
    sc = scenes.config_hull(f=1.5e3, n_angles=4, **SMALL)
    _banner("hull SELF mask bit-exact (924 triangles, LBVH)")
    runner.load_scene(sc)
    s = oracle.Scene()
    s.set_initial_conditions(sc["c"], sc["f"], sc["atten"])
    s.load_geometry(sc["targets"][0][0], 0, sc["targets"][0][1])
    s.load_source_points(sc["sources"])
    s.load_field_points(sc["points"])
    s.render()
    c = s.cells(0, 0)
    fr = s.frames(0, 0)
    n = fr[c["facet"], 0:3].astype(np.float32)
    act = c["area"] > 1e-6
    ref = maskref.self_mask(s, c["pe"][act], n[act], c["facet"][act])
    na = int(act.sum())
    got = api.mask_bits(2, na, na, ta=0, ia=0)
    print("SELF visible", ref.sum(), "of", ref.size)
    assert ref.sum() > 0 and np.array_equal(got, ref)
    api.TearDownCuda()


def test_hull_plain_and_transfer_vs_oracle(api):
    sc = scenes.config_hull(f=2e3, n_angles=16, **SMALL)
    _banner("hull 2 kHz, 16 incidences, 2 bounces, attenuation: plain and transfer forms vs oracle")
    plain = _run(api, sc, 0)
    mixed = _run(api, sc, 1)
    for p in (0, 5, 11):
        ref = oracle.run_scene(sc, plane_source=p)
        b = np.concatenate([ref["points"], ref["targets"][0]])
        ep, em = rel_l2(_flat(plain, p), b), rel_l2(_flat(mixed, p), b)
        print("plane", p, "plain", ep, "transfer", em)
        assert ep < 2e-4 and em < 1e-3


def test_hull_full_size_transfer_matches_plain(api):
    sc = scenes.config_hull(n_angles=64)  # 8 kHz, lambda/7.5: 63,764 fragments
    _banner("hull 8 kHz full mesh, 64 incidences: transfer form vs plain form")
    plain = _run(api, sc, 0)
    mixed = _run(api, sc, 1)
    a = np.concatenate([plain["points"].ravel(), plain["targets"][0].ravel()])
    b = np.concatenate([mixed["points"].ravel(), mixed["targets"][0].ravel()])
    e = rel_l2(b, a)
    print("transfer vs plain rel-L2", e, "points only", rel_l2(mixed["points"].ravel(), plain["points"].ravel()))
    assert np.all(np.isfinite(a)) and e < 1e-3
