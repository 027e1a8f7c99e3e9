"""The allocation-free field-point projection against ProjectTargetToFieldPoints.

sr_project_target_to_field_points_noalloc sends its split-K partial sums to a
fixed device array instead of the library's scratch and allocates nothing.
Each case runs one scene twice in one context: once through the allocating
entry point, once (after resetting the pressures and repeating the same
projections) through the allocation-free one, whose output must equal the
first. Three calls in a row, each on a smaller input than the one before,
so a read of the previous call's leftover partial sums would show: a case whose
partials exceed the fixed array (fewer emitter splits there, so equal to
FP32 rounding), the 64-incidence trihedral (bit for bit) and the single-incidence
plate (bit for bit).
"""
import numpy as np
import pytest

import scenes
from conftest import rel_l2


def _points(api, sc):
    nq = len(sc["points"])
    return np.stack([api.field_points(nq, p) for p in range(api.num_planes())])


def _case(api, sc, exact):
    import runner
    runner.load_scene(sc)
    runner.project_scene(sc)
    want = _points(api, sc)
    api.reset_pressures()
    api.project_source_points_to_objects(sc.get("project_field_objects", False))
    if sc.get("bounces", 0):
        api.project_target_to_target_objects(sc["bounces"])
    api.device_synchronize()
    before = api.device_allocations()
    api.project_target_to_field_points_noalloc()
    got = _points(api, sc)
    assert api.device_allocations() == before, "the allocation-free call allocated device memory"
    api.TearDownCuda()
    if exact:
        assert np.array_equal(got, want), rel_l2(got, want)
    else:
        assert rel_l2(got, want) < 1e-6, rel_l2(got, want)
    assert np.abs(want).max() > 0


@pytest.mark.gpu
def test_noalloc_field_points_match_allocating(api):
    # 1) partial sums beyond the fixed array: 4,096 field points x 64 incidences
    sc = scenes.config_trihedral(f=10e3, n_side=8, bounces=1)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(4096, 3))
    d = np.abs(d) / np.linalg.norm(d, axis=1, keepdims=True)
    sc = dict(sc, points=(15.0 * d).astype(np.float32))
    _case(api, sc, exact=False)
    # 2) the 64-incidence trihedral, monostatic points: fits, same splits
    _case(api, scenes.config_trihedral(f=10e3, n_side=8, bounces=2), exact=True)
    # 3) a smaller input: the plate, one incidence, 61 points (the vector-unit gather)
    _case(api, scenes.config_plate(n_points=61), exact=True)


@pytest.mark.gpu
def test_noalloc_rejects_before_render(api):
    with pytest.raises(api.SonarRenderError):
        api.project_target_to_field_points_noalloc()
