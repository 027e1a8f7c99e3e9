"""C-ABI checks that need no GPU: every declared symbol is exported, and the
error paths the reference turned into std::terminate return status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sonar_render.h")

REFERENCE_SYMBOLS = [  # src/python_interface.cpp:14-107
    "set_initial_conditions", "load_geometry", "load_field_points", "load_source_points", "render_cuda",
    "ProjectSrcPointsToObjects", "ProjectTargetToFieldObjects", "ProjectTargetToFieldPoints",
    "ProjectTargetToTargetObjects", "GetFieldPointPressures", "RenderOpenGL", "TearDownCuda",
    "SoundVisualisationInit",
]


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\([^;{]*\)\s*;", text)))


def test_header_declares_reference_symbols():
    decl = _declared()
    for s in REFERENCE_SYMBOLS:
        assert s in decl


def test_library_exports_every_declared_symbol():
    import sonar_api
    lib = ctypes.CDLL(sonar_api.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_error_paths_return_status(api):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    rc = api.cpp_lib.sr_load_geometry(tri.ctypes.data_as(api._fp), 1, 7, 0.1)
    assert rc == -2 and "invalid object type" in api.cpp_lib.sr_last_error().decode()
    assert api.cpp_lib.sr_project_target_to_target_objects(1) == -3
    assert api.cpp_lib.sr_project_src_points_to_objects(0) == -3
    out = np.zeros(4, np.float64)
    assert api.cpp_lib.sr_get_field_point_pressures(out.ctypes.data_as(api._dp), 2) == -3
    assert api.cpp_lib.sr_load_geometry(tri.ctypes.data_as(api._fp), 1, 0, 0.0) == -2
    # the reference's void entry point with a bad type must not abort the process
    api.cpp_lib.load_geometry.argtypes = [api._fp, ctypes.c_int, ctypes.c_int, ctypes.c_float]
    api.cpp_lib.load_geometry(tri.ctypes.data_as(api._fp), 1, 9, 0.1)
    assert api.cpp_lib.sr_last_status() == -2
    with pytest.raises(api.SonarRenderError):
        api.load_stl_mesh_to_cuda(tri[None], 5, 0.1)
    # the render needs a rendered scene (status -3), not a GPU call
    with pytest.raises(api.SonarRenderError, match="-3"):
        api.render_openGL(8, 8, [0.0, 0, 3, 0, 0, 0, 0, 1, 0])
    with pytest.raises(api.SonarRenderError):
        api.sound_visualisation_init(-90.0, -20.0, False, False)  # max below min


def test_source_objects_are_counted(api):
    tri = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    api.load_stl_mesh_to_cuda(tri, 1, 0.1)
    assert api.object_count(1) == 1 and api.object_count(0) == 0


def test_status_is_per_thread():
    """Two threads, each on its own context, fail concurrently with different
    errors; each reads back its own status and message (sr_last_error)."""
    import threading
    import sonar_api as api
    lib = api.cpp_lib
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    bad = []
    barrier = threading.Barrier(2)

    def worker(kind):
        h = lib.sr_context_create()
        assert h > 0 and lib.sr_context_select(h) == 0
        barrier.wait()
        for _ in range(2000):
            if kind == 0:
                rc = lib.sr_load_geometry(tri.ctypes.data_as(api._fp), 1, 7, 0.1)
                want, rcw = "invalid object type", -2
            else:
                rc = lib.sr_project_src_points_to_objects(0)
                want, rcw = "render_cuda", -3
            msg = lib.sr_last_error().decode()
            if rc != rcw or lib.sr_last_status() != rcw or want not in msg:
                bad.append((kind, rc, msg))
                break
        lib.sr_context_select(0)
        lib.sr_context_destroy(h)

    ts = [threading.Thread(target=worker, args=(k,)) for k in (0, 1)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not bad, bad


def test_fragment_probe_shapes(api):
    (nx, ny, nneg), area, pe, pr = api.fragment_facet(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), 0.1)
    assert area.shape == (ny, nx) and pe.shape == (nx * ny, 3) and pr.shape == (nx * ny, 3)
    assert 0 <= nneg <= nx
    assert abs(area.sum() - 0.5) < 1e-5


def test_contexts_without_gpu():
    """Context handles (no device work until render): each has its own scene state,
    the calling thread's current one receives the calls, and misuse returns errors."""
    import threading

    import sonar_api as api
    assert api.context_current() == 0
    h1, h2 = api.context_create(), api.context_create()
    assert h1 > 0 and h2 > 0 and h1 != h2
    try:
        api.context_select(h1)
        assert api.context_current() == h1
        api.set_batched_sources(True)
        api.load_points_to_cuda(np.zeros((3, 3), np.float32), isSource=True)
        api.context_select(h2)
        api.load_points_to_cuda(np.zeros((5, 3), np.float32), isSource=True)
        # another thread starts on the default context, whatever this one selected
        seen = []
        t = threading.Thread(target=lambda: seen.append(api.context_current()))
        t.start()
        t.join()
        assert seen == [0]
        with pytest.raises(api.SonarRenderError):
            api.context_destroy(h2)  # current on this thread
        with pytest.raises(api.SonarRenderError):
            api.context_destroy(0)
        with pytest.raises(api.SonarRenderError):
            api.context_select(10 ** 6)
        assert api.context_current() == h2  # a failed select keeps the current context
    finally:
        api.context_select(0)
        api.context_destroy(h1)
        api.context_destroy(h2)
    with pytest.raises(api.SonarRenderError):
        api.context_select(h1)  # destroyed
    assert api.context_current() == 0


def test_form_settings_checked_and_reset(api):
    """The settings that choose between bit-identical ways of computing a form
    take their documented values only (status -2 otherwise, the setting kept),
    and TearDownCuda puts them back to their defaults. No GPU call is made."""
    lib = api.cpp_lib
    cases = [("sr_set_transfer_symmetric", "sr_transfer_symmetric", (0, 1, 2), (-1, 3), 2),
             ("sr_set_inline_visibility", "sr_inline_visibility", (0, 1), (-1, 2), 1),
             ("sr_set_reduce_in_product", "sr_reduce_in_product", (0, 1), (-1, 2), 0)]
    for setter, getter, good, bad, default in cases:
        set_, get = getattr(lib, setter), getattr(lib, getter)
        assert get() == default, setter
        for v in good:
            assert set_(v) == 0 and get() == v, (setter, v)
        for v in bad:
            assert set_(v) == -2 and get() == good[-1], (setter, v)
            assert lib.sr_last_error().decode(), setter
    api.TearDownCuda()
    for setter, getter, good, bad, default in cases:
        assert getattr(lib, getter)() == default, getter
