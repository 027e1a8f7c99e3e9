"""GPU: the sphere and trihedral workloads at their full sizes.

* sphere r = 1 m, icosphere of 5,120 triangles, 10 kHz, lambda/7.5 (59,680
  fragments), 360 bistatic points on a 30 m ring and a 256^2-fragment field
  plane: point->fragment and fragment->point visibility words bit-exact against
  the oracle's definition (src/OptiX.cu:81-93,163-175), seeded receiver blocks of
  the fragment->field-plane words bit-exact (OptiX.cu:321-333), the pressures
  against the FP64 oracle, and seeded field-plane cells against 0.9 times the
  oracle's fragment->point sum at their points (no cell is within a fragment
  size of the sphere, so the r < delta branch never applies);
* trihedral 1 m, 20 kHz, lambda/7.5 (15,768 fragments), 64 incidences, 3
  bounces: seeded rows of the SELF words bit-exact, the plain and FP16 transfer
  forms against the FP64 oracle on 2 seeded incidences.
"""
import numpy as np
import pytest

import maskref
import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def _tables(s, t, i):
    c = s.cells(t, i)
    fr = s.frames(t, i)
    n = fr[c["facet"], 0:3].astype(np.float32)
    act = c["area"] > 1e-6
    return c, n, act


def _block_rows(api, kind, n_emit, blocks, **kw):
    """Rows (receivers) of the chosen 64-receiver blocks as bool [len(blocks)*64, n_emit]."""
    w = api.mask_words(kind, **kw).reshape(-1, n_emit)[list(blocks)]
    bits = np.unpackbits(w.view(np.uint8).reshape(len(blocks), n_emit, 8), axis=2, bitorder="little")
    return bits.transpose(0, 2, 1).reshape(-1, n_emit).astype(bool)


def _oracle_scene(sc):
    s = oracle.Scene()
    s.set_initial_conditions(sc["c"], sc["f"], sc.get("atten", 0.0))
    for tris, d in sc["targets"]:
        s.load_geometry(tris, 0, d)
    for tris, d in sc.get("fields", []):
        s.load_geometry(tris, 2, d)
    s.load_source_points(sc["sources"])
    s.load_field_points(sc["points"])
    s.render()
    return s


def test_sphere_full_size(api):
    sc = scenes.config_sphere()  # 10 kHz, 5,120 triangles, 360 points, 256^2 field plane
    _banner("sphere 10 kHz full size: masks, pressures, field plane")
    runner.load_scene(sc)
    runner.project_scene(sc)
    out = runner.read_results(sc)
    s = _oracle_scene(sc)
    c, n, act = _tables(s, 0, 0)
    cf, nf, actf = _tables(s, 2, 0)
    na, nfa, nq = int(act.sum()), int(actf.sum()), len(sc["points"])
    print("fragments", na, "field fragments", nfa, flush=True)
    assert na == 59680
    # point -> fragment over every cell, fragment -> point: whole sets
    ref = maskref.p2f(s, sc["sources"], c["pr"], n, True, cell_tri=c["facet"])
    assert np.array_equal(api.mask_bits(0, 1, len(c["area"]), tb=0, ib=0), ref)
    ref = maskref.f2p(s, c["pe"][act], n[act], sc["points"], em_tri=c["facet"][act])
    got = api.mask_bits(1, na, nq, ta=0, ia=0)
    print("F2P visible", ref.sum(), "of", ref.size, flush=True)
    assert ref.sum() > 0 and np.array_equal(got, ref)
    # fragment -> field plane: seeded 64-receiver blocks
    rng = np.random.default_rng(2)
    nrb = (nfa + 63) // 64
    blocks = np.sort(rng.choice(nrb, 6, replace=False))
    got = _block_rows(api, 3, na, blocks, ta=0, ia=0, tb=2, ib=0)
    rows = np.concatenate([np.arange(b * 64, min(nfa, b * 64 + 64)) for b in blocks])
    ref = maskref.f2f(s, c["pe"][act], n[act], cf["pe"][actf][rows], None, rc_field=True, em_tri=c["facet"][act])
    print("NO_RESP rows visible", ref.sum(), "of", ref.size, flush=True)
    assert ref.sum() > 0 and np.array_equal(got[:len(rows)], ref)
    s.close()
    # pressures: points and the sphere's cells against the oracle
    ref = oracle.run_scene(dict(sc, fields=[], project_field_objects=False))
    e_pts, e_tgt = rel_l2(out["points"][0], ref["points"]), rel_l2(out["targets"][0][0], ref["targets"][0])
    print("points", e_pts, "sphere cells", e_tgt, flush=True)
    assert e_pts < 2e-4 and e_tgt < 2e-4
    # field plane: seeded cells at least two fragment sizes from the sphere
    d = sc["fields"][0][1]
    idx = np.where(actf)[0]
    radius = np.linalg.norm(cf["pe"][idx].astype(np.float64), axis=1)
    far = idx[radius > 1.0 + 2 * d]
    pick = np.sort(rng.choice(far, 256, replace=False))
    ref = oracle.run_scene(dict(sc, fields=[], project_field_objects=False, points=cf["pe"][pick]))
    got = out["fields"][0][0][pick]
    e = rel_l2(got, 0.9 * ref["points"])
    print("field plane cells vs 0.9 x oracle fragment->point", e, flush=True)
    assert e < 2e-4


def test_trihedral_20khz_full_size(api):
    sc = scenes.config_trihedral()  # 20 kHz, 64 incidences, 3 bounces
    _banner("trihedral 20 kHz full size: SELF rows, plain and transfer vs the oracle")
    runner.load_scene(sc)
    s = _oracle_scene(sc)
    c, n, act = _tables(s, 0, 0)
    na = int(act.sum())
    assert na == 15768
    rng = np.random.default_rng(3)
    blocks = np.sort(rng.choice((na + 63) // 64, 3, replace=False))
    got = _block_rows(api, 2, na, blocks, ta=0, ia=0)
    rows = np.concatenate([np.arange(b * 64, min(na, b * 64 + 64)) for b in blocks])
    p, nn, fa = c["pe"][act], n[act], c["facet"][act]
    # rows of the SELF set: the ray leaves the lower facet of each pair
    fwd = maskref.f2f(s, p, nn, p[rows], nn[rows], em_tri=fa, rc_tri=fa[rows])            # from the emitter
    rev = maskref.f2f(s, p[rows], nn[rows], p, nn, em_tri=fa[rows], rc_tri=fa).T          # from the receiver
    ref = np.where(fa[None, :] < fa[rows][:, None], fwd, np.where(fa[None, :] > fa[rows][:, None], rev, False))
    print("SELF rows visible", ref.sum(), "of", ref.size, flush=True)
    assert ref.sum() > 0 and np.array_equal(got[:len(rows)], ref)
    s.close()
    api.TearDownCuda()
    outs = {m: runner.run_scene(sc, transfer_mode=m) for m in (0, 1)}
    for p in (5, 50):
        ref = oracle.run_scene(sc, plane_source=p)
        b = np.concatenate([ref["points"], ref["targets"][0]])
        e = [rel_l2(np.concatenate([outs[m]["points"][p], outs[m]["targets"][0][p]]), b) for m in (0, 1)]
        print("incidence", p, "plain", e[0], "transfer", e[1], flush=True)
        assert e[0] < 2e-4 and e[1] < 1e-3
