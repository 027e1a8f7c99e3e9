"""GPU: two scene contexts used in turn from one thread. Each keeps its own
geometry, sources, settings, device buffers and streams; results equal the same
scenes run alone on the default context, bit for bit."""
import numpy as np
import pytest

import runner
import scenes

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(f"[gpu] {msg}", flush=True)


def test_two_contexts_interleaved(api):
    plate = scenes.config_plate(f=5e3, n_points=61)
    tri = scenes.config_trihedral(f=5e3, n_side=2, bounces=2)
    _banner("contexts: plate (plain) and trihedral (transfer form), interleaved")
    ref_plate = runner.run_scene(plate)
    ref_tri = runner.run_scene(tri, transfer_mode=1)
    a, b = api.context_create(), api.context_create()
    try:
        api.context_select(a)
        runner.load_scene(plate)
        api.context_select(b)
        runner.load_scene(tri, transfer_mode=1)
        for _ in range(2):  # two scenes each, the contexts' work queued in turn
            api.context_select(a)
            api.reset_pressures()
            runner.project_scene(plate)
            api.context_select(b)
            api.reset_pressures()
            runner.project_scene(tri)
        api.context_select(a)
        out_plate = runner.read_results(plate)
        api.TearDownCuda()
        api.context_select(b)
        out_tri = runner.read_results(tri)
        api.TearDownCuda()
    finally:
        api.context_select(0)
        api.context_destroy(a)
        api.context_destroy(b)
    assert np.array_equal(out_plate["points"], ref_plate["points"])
    assert np.array_equal(out_tri["points"], ref_tri["points"])
    assert np.array_equal(out_tri["targets"][0], ref_tri["targets"][0])
    # the default context is untouched and still works
    again = runner.run_scene(plate)
    assert np.array_equal(again["points"], ref_plate["points"])
