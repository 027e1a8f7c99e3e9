"""Source objects (type 1): piston-like emitters with one coherent amplitude per
object (README.md:2), stored but never used by the reference (src/Model.cpp:33-35).

Here a source object's fragments radiate through the reference's
fragment->fragment sum without reciprocity (src/OptiX.cu:192-406, FACET_NO_RESP
form) into the total pressure P of every target (and of every field object when
ProjectSrcPointsToObjects(true)). The CPU test pins the oracle against a dense
numpy sum over the piston's cells; the GPU test compares the library with the
oracle.
"""
import math

import numpy as np
import pytest

import oracle
import scenes
from conftest import rel_l2


def test_oracle_source_object_dense_sum_and_linearity():
    sc = scenes.config_piston(f=5e3, n_points=31)
    s = oracle.Scene()
    s.set_initial_conditions(sc["c"], sc["f"], 0.0)
    tris_t, dt = sc["targets"][0]
    tris_s, ds = sc["source_objects"][0]
    s.load_geometry(tris_t, 0, dt)
    s.load_geometry(tris_s, 1, ds)
    s.load_field_points(sc["points"])
    s.render()
    s.set_source_object_amplitude(0, 0.5 - 0.25j)
    s.project_src(False)
    got = s.pressure(0, 0)
    # dense restatement: receivers = active target cells at their fragment points
    ct = s.cells(0, 0)
    cs_ = s.cells(1, 0)
    ft, fs = s.frames(0, 0), s.frames(1, 0)
    k = complex(s.kf(), 0.0)
    ea = np.where(cs_["area"] > 1e-6)[0]
    ra = np.where(ct["area"] > 1e-6)[0]
    want = np.zeros(len(ct["area"]), complex)
    for i in ea:
        pe = cs_["pe"][i].astype(np.float64)
        ne = fs[cs_["facet"][i], 0:3].astype(np.float64)
        v = ct["pe"][ra].astype(np.float64) - pe
        r = np.linalg.norm(v, axis=1)
        u = v / r[:, None]
        c1 = u @ ne
        c2 = -(u * ft[ct["facet"][ra], 0:3]).sum(1)
        ok = (c1 >= 1e-6) & (c2 >= 1e-6)
        G = np.where(r < ds, 1j * k / (4 * math.pi),
                     -np.exp(1j * k * r) / (4 * math.pi) * (1j * k / r) + 1.0 / (r * r))
        want[ra] += np.where(ok, cs_["area"][i] * c1 * c2 * 0.9 * G * (0.5 - 0.25j), 0)
    # nothing occludes between a piston and a flat plate; FP32 geometry in the oracle
    assert rel_l2(got, want) < 1e-5, rel_l2(got, want)
    assert np.count_nonzero(got) == np.count_nonzero(want) > 0
    s.close()


def _cells_of(s, t, i):
    c = s.cells(t, i)
    return c


@pytest.mark.gpu
def test_source_object_matches_oracle(api):
    import runner
    for fs in (False, True):
        sc = scenes.config_piston(f=10e3, n_points=61, field_surface=fs)
        ref = oracle.run_scene(sc)
        out = runner.run_scene(sc)
        assert rel_l2(out["targets"][0][0], ref["targets"][0]) < 2e-4
        assert rel_l2(out["points"][0], ref["points"]) < 2e-4
        if fs:
            assert rel_l2(out["fields"][0][0], ref["fields"][0]) < 2e-4
    # amplitude scales the result; batched source points are refused with source objects
    sc = scenes.config_piston(f=10e3, n_points=61)
    runner.load_scene(sc)
    api.set_source_object_amplitude(0, 2.0 - 1.0j)
    runner.project_scene(sc)
    res = runner.read_results(sc)
    base = runner.run_scene(sc)
    assert rel_l2(res["points"][0], (2.0 - 1.0j) * base["points"][0]) < 1e-5
    api.TearDownCuda()
    api.set_batched_sources(True)
    api.load_points_to_cuda(np.array([[0, 0, 5.0]], np.float32), isSource=True)
    api.load_stl_mesh_to_cuda(sc["targets"][0][0], 0, sc["targets"][0][1])
    api.load_stl_mesh_to_cuda(sc["source_objects"][0][0], 1, sc["source_objects"][0][1])
    with pytest.raises(api.SonarRenderError):
        api.render_cuda()
