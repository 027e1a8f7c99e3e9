"""ctypes binding of libsonar_render.so.

The first group of functions keeps the names, arguments and behaviour of the
reference's Python wrapper (run_scripts/modules/api.py:6-112) so scene scripts
written for the reference run unchanged. Two reference slips are fixed and
documented: ``set_initial_conditions`` passes the fourth (density) argument the
C function declares (api.py:11-14 passes three), and
``project_target_to_field_points`` calls ProjectTargetToFieldPoints
(api.py:96-98 calls ProjectTargetToFieldObjects).

The second group exposes the library's status-returning extensions; every one
raises :class:`SonarRenderError` with the library's message on failure.

The library is loaded from ``lib/libsonar_render.so`` beside this file (or the
path in ``SONAR_RENDER_LIB``). A missing library raises at import: there is no
fallback implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SONAR_RENDER_LIB", os.path.join(_HERE, "lib", "libsonar_render.so"))

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libsonar_render.so not built: {LIB_PATH} (run __graft_entry__.build())")

cpp_lib = ctypes.CDLL(LIB_PATH)

_f = ctypes.c_float
_i = ctypes.c_int
_l = ctypes.c_long
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u64p = ctypes.POINTER(ctypes.c_ulonglong)


def _sig(name, args, res=ctypes.c_int):
    fn = getattr(cpp_lib, name)
    fn.argtypes = args
    fn.restype = res
    return fn


_sig("sr_last_error", [], ctypes.c_char_p)
_sig("sr_last_status", [])
for _n, _a in [
    ("sr_set_initial_conditions", [_f, _f, _f, _f]),
    ("sr_load_geometry", [_fp, _i, _i, _f]),
    ("sr_load_field_points", [_fp, _i]),
    ("sr_load_source_points", [_fp, _i]),
    ("sr_render", []),
    ("sr_project_src_points_to_objects", [_i]),
    ("sr_project_target_to_field_objects", []),
    ("sr_project_target_to_field_points", []),
    ("sr_project_target_to_target_objects", [_i]),
    ("sr_get_field_point_pressures", [_dp, _i]),
    ("sr_get_field_point_pressures_c64", [_fp, _i]),
    ("sr_teardown", []),
    ("sr_set_batched_sources", [_i]),
    ("sr_num_planes", []),
    ("sr_get_field_point_pressures_plane", [_i, _dp, _i]),
    ("sr_set_source_amplitude", [_i, _f, _f]),
    ("sr_prepare_resp", [_i]),
    ("sr_project_resp", [_i, _i]),
    ("sr_object_count", [_i]),
    ("sr_object_active_count", [_i, _i]),
    ("sr_object_active_cells", [_i, _i, _ip, _i]),
    ("sr_get_object_pressures", [_i, _i, _i, _i, _dp, _l]),
    ("sr_get_mask", [_i, _i, _i, _i, _i, _u64p, _l]),
    ("sr_vis_params", [_fp]),
]:
    _sig(_n, _a)
_sig("sr_object_cell_count", [_i, _i], _l)
_sig("sr_mask_words", [_i, _i, _i, _i, _i], _l)
_sig("sr_fragment_facet", [_fp, _f, _ip, _fp, _fp, _fp, _l], _l)
_sig("sr_last_call_ms", [], ctypes.c_float)
_u8p = ctypes.POINTER(ctypes.c_ubyte)
_sig("sr_sound_visualisation_init", [_f, _f, _i, _i])
_sig("sr_set_source_object_amplitude", [_i, _f, _f])
_sig("sr_project_target_to_field_points_noalloc", [])
_sig("sr_device_allocations", [], _l)
_sig("sr_batch_set_near", [_i, _f, _f])
_sig("sr_render_png", [ctypes.c_char_p, _i, _i, _fp, _i])
_sig("sr_render_rgb", [_i, _i, _i, _fp, _u8p])
_sig("sr_render_facets_rgb", [_fp, _i, _f, _dp, _l, _i, _f, _f, _i, _i, _fp, _u8p])
_sig("sr_write_png", [ctypes.c_char_p, _i, _i, _u8p])


class SonarRenderError(RuntimeError):
    pass


def _check(rc, what):
    if rc < 0:
        msg = cpp_lib.sr_last_error().decode("utf-8", "replace")
        raise SonarRenderError(f"{what} failed ({rc}): {msg}")
    return rc


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ptr(a, t=_fp):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------
# Reference-compatible wrapper (run_scripts/modules/api.py)

def set_initial_conditions(cp, frequency, attenuation, density=0.0):
    """Sets the medium: sound speed, frequency, attenuation (used as Im k)."""
    _check(cpp_lib.sr_set_initial_conditions(cp, frequency, attenuation, density), "set_initial_conditions")


def load_points_to_cuda(points, isSource=False):
    """Loads source or field points (N x 3)."""
    pts = _f32(points).reshape(-1, 3)
    fn = cpp_lib.sr_load_source_points if isSource else cpp_lib.sr_load_field_points
    _check(fn(_ptr(pts), len(pts)), "load points")


def load_stl_mesh_to_cuda(stl_mesh, object_type, resolution=2e-3):
    """Loads a mesh (anything with ``.vectors`` of shape (N,3,3), or such an array)."""
    vec = getattr(stl_mesh, "vectors", stl_mesh)
    v = _f32(vec).reshape(-1, 9)
    _check(cpp_lib.sr_load_geometry(_ptr(v), len(v), int(object_type), resolution), "load_geometry")


def render_cuda():
    """Fragments the facets, builds the LBVH and uploads the scene."""
    _check(cpp_lib.sr_render(), "render_cuda")


def GetFieldPoints(NumFieldPnts):
    """Field-point pressures as an (N, 2) float64 array of (real, imag)."""
    out = np.zeros((NumFieldPnts, 2), dtype=np.float64)
    _check(cpp_lib.sr_get_field_point_pressures(_ptr(out, _dp), NumFieldPnts), "GetFieldPointPressures")
    return out


def render_openGL(width, height, viewSettings, filename="output"):
    """Writes ``filename + ".png"``: the scene's surfaces seen from viewSettings
    (eye, look-at point, up vector), as api.py:70-79 (rendered on the host)."""
    view = _f32(viewSettings).reshape(9)
    _check(cpp_lib.sr_render_png((filename + ".png").encode("utf-8"), int(width), int(height), _ptr(view), 0),
           "RenderOpenGL")


def TearDownCuda():
    _check(cpp_lib.sr_teardown(), "TearDownCuda")


def project_source_points_to_objects(project_to_field_surface=False):
    _check(cpp_lib.sr_project_src_points_to_objects(1 if project_to_field_surface else 0), "ProjectSrcPointsToObjects")


def project_target_to_field_objects():
    _check(cpp_lib.sr_project_target_to_field_objects(), "ProjectTargetToFieldObjects")


def project_target_to_field_points():
    _check(cpp_lib.sr_project_target_to_field_points(), "ProjectTargetToFieldPoints")


def project_target_to_field_points_noalloc():
    """ProjectTargetToFieldPoints without any device allocation (fixed scratch)."""
    _check(cpp_lib.sr_project_target_to_field_points_noalloc(), "ProjectTargetToFieldPoints (allocation-free)")


def device_allocations():
    """Device allocations the scene model has made so far in this process."""
    return int(cpp_lib.sr_device_allocations())


def project_target_to_target_objects(number_reflections):
    _check(cpp_lib.sr_project_target_to_target_objects(int(number_reflections)), "ProjectTargetToTargetObjects")


def sound_visualisation_init(max_db, min_db, render_phase_target, render_phase_field):
    """Colour range in dB and phase flags of the render (python_interface.cpp:103-107)."""
    _check(cpp_lib.sr_sound_visualisation_init(float(max_db), float(min_db), 1 if render_phase_target else 0,
                                               1 if render_phase_field else 0), "SoundVisualisationInit")


def render_rgb(width, height, viewSettings, plane=0):
    """The render as a (height, width, 3) uint8 array, top row first."""
    view = _f32(viewSettings).reshape(9)
    out = np.zeros((height, width, 3), np.uint8)
    _check(cpp_lib.sr_render_rgb(int(plane), int(width), int(height), _ptr(view),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))), "render_rgb")
    return out


def render_facets_rgb(tris, delta, pressures, width, height, viewSettings, phase=False, min_db=-90.0, max_db=-20.0):
    """Host-only render of one object (triangles, fragment edge, complex per-cell pressures)."""
    v = _f32(tris).reshape(-1, 9)
    p = np.ascontiguousarray(np.stack([np.real(pressures), np.imag(pressures)], 1).astype(np.float64))
    view = _f32(viewSettings).reshape(9)
    out = np.zeros((height, width, 3), np.uint8)
    _check(cpp_lib.sr_render_facets_rgb(_ptr(v), len(v), float(delta), _ptr(p, _dp), len(p), 1 if phase else 0,
                                        float(min_db), float(max_db), int(width), int(height), _ptr(view),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))), "render_facets_rgb")
    return out


def write_png(path, rgb):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w = rgb.shape[:2]
    _check(cpp_lib.sr_write_png(path.encode("utf-8"), w, h, rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))),
           "write_png")


# ---------------------------------------------------------------------------
# Extensions

def set_batched_sources(on=True):
    _check(cpp_lib.sr_set_batched_sources(1 if on else 0), "set_batched_sources")


def num_planes():
    return cpp_lib.sr_num_planes()


def field_points(n, plane=0):
    """complex128 field-point pressures of one incidence plane."""
    out = np.zeros((n, 2), dtype=np.float64)
    _check(cpp_lib.sr_get_field_point_pressures_plane(plane, _ptr(out, _dp), n), "field_points")
    return out[:, 0] + 1j * out[:, 1]


def field_points_c64(n):
    out = np.zeros((n, 2), dtype=np.float32)
    _check(cpp_lib.sr_get_field_point_pressures_c64(_ptr(out), n), "field_points_c64")
    return out[:, 0] + 1j * out[:, 1]


def set_source_amplitude(i, value):
    _check(cpp_lib.sr_set_source_amplitude(i, float(np.real(value)), float(np.imag(value))), "set_source_amplitude")


def set_source_object_amplitude(i, value):
    _check(cpp_lib.sr_set_source_object_amplitude(i, float(np.real(value)), float(np.imag(value))),
           "set_source_object_amplitude")


def prepare_resp(target):
    _check(cpp_lib.sr_prepare_resp(target), "prepare_resp")


def project_resp(a, b):
    _check(cpp_lib.sr_project_resp(a, b), "project_resp")


def object_count(object_type):
    return cpp_lib.sr_object_count(object_type)


def cell_count(object_type, index):
    return _check(cpp_lib.sr_object_cell_count(object_type, index), "cell_count")


def active_cells(object_type, index):
    n = _check(cpp_lib.sr_object_active_count(object_type, index), "active_count")
    out = np.zeros(max(n, 1), dtype=np.int32)
    _check(cpp_lib.sr_object_active_cells(object_type, index, _ptr(out, _ip), n), "active_cells")
    return out[:n]


def object_pressures(object_type, index, which=0, plane=0):
    """complex128 per-cell pressure (which: 0 P, 1 P_in, 2 P_out)."""
    n = cell_count(object_type, index)
    out = np.zeros((max(n, 1), 2), dtype=np.float64)
    _check(cpp_lib.sr_get_object_pressures(object_type, index, which, plane, _ptr(out, _dp), n), "object_pressures")
    return out[:n, 0] + 1j * out[:n, 1]


def mask_words(kind, ta=0, ia=0, tb=0, ib=0):
    """Visibility words [rblock][emitter] as uint64."""
    n = _check(cpp_lib.sr_mask_words(kind, ta, ia, tb, ib), "mask_words")
    out = np.zeros(max(n, 1), dtype=np.uint64)
    _check(cpp_lib.sr_get_mask(kind, ta, ia, tb, ib, _ptr(out, _u64p), n), "get_mask")
    return out[:n]


def mask_bits(kind, n_emit, n_recv, ta=0, ia=0, tb=0, ib=0):
    """Visibility as a bool array [receiver, emitter]."""
    w = mask_words(kind, ta, ia, tb, ib).reshape(-1, n_emit)
    bits = np.unpackbits(w.view(np.uint8).reshape(w.shape[0], n_emit, 8), axis=2, bitorder="little")
    bits = bits.transpose(0, 2, 1).reshape(-1, n_emit)
    return bits[:n_recv].astype(bool)


def vis_params():
    out = np.zeros(5, dtype=np.float32)
    _check(cpp_lib.sr_vis_params(_ptr(out)), "vis_params")
    return out


def last_call_ms():
    return float(cpp_lib.sr_last_call_ms())


def fragment_facet(tri, delta):
    """Host fragmenter probe: (grid (nX,nY,nNeg), area[nY,nX], emitter pts, receiver pts)."""
    v = _f32(tri).reshape(9)
    g = np.zeros(3, dtype=np.int32)
    n = _check(cpp_lib.sr_fragment_facet(_ptr(v), delta, _ptr(g, _ip), None, None, None, 0), "fragment_facet")
    area = np.zeros(max(n, 1), np.float32)
    pe = np.zeros((max(n, 1), 3), np.float32)
    pr = np.zeros((max(n, 1), 3), np.float32)
    _check(cpp_lib.sr_fragment_facet(_ptr(v), delta, _ptr(g, _ip), _ptr(area), _ptr(pe), _ptr(pr), n), "fragment_facet")
    nx, ny = int(g[0]), int(g[1])
    return (nx, ny, int(g[2])), area[:n].reshape(ny, nx), pe[:n], pr[:n]


_sig("sr_reset_pressures", [])
_sig("sr_clear_vis_cache", [])
_sig("sr_profile", [_i])
_sig("sr_kernel_stats", [_i, _dp, ctypes.POINTER(ctypes.c_long), _dp, _dp, _dp])

KERNEL_FAMILIES = {"p2f": 0, "vis": 1, "f2f": 2, "f2f_field": 3, "f2p": 4, "reduce": 5}


def reset_pressures():
    _check(cpp_lib.sr_reset_pressures(), "reset_pressures")


def clear_vis_cache():
    _check(cpp_lib.sr_clear_vis_cache(), "clear_vis_cache")


def profile(on=True):
    _check(cpp_lib.sr_profile(1 if on else 0), "profile")


def kernel_stats(family):
    """dict(ms, launches, pairs, flops, bytes) summed over the recorded launches of one kernel family."""
    ms, pairs, flops, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    n = ctypes.c_long()
    _check(cpp_lib.sr_kernel_stats(KERNEL_FAMILIES[family], ctypes.byref(ms), ctypes.byref(n), ctypes.byref(pairs),
                                   ctypes.byref(flops), ctypes.byref(by)), "kernel_stats")
    return dict(ms=ms.value, launches=n.value, pairs=pairs.value, flops=flops.value, bytes=by.value)


_sig("sr_time_transfer_gemm", [_i, ctypes.POINTER(ctypes.c_double)])


def time_transfer_gemm(reps=20):
    """Mean ms per launch of the last transfer-form bounce product, replayed back to back."""
    ms = ctypes.c_double()
    _check(cpp_lib.sr_time_transfer_gemm(int(reps), ctypes.byref(ms)), "time_transfer_gemm")
    return ms.value


_sig("sr_measure_stream_read", [_l, _i, ctypes.POINTER(ctypes.c_double)])


def measure_stream_read(mib=1024, reps=20):
    """Attainable HBM read bandwidth (GB/s): a plain streaming read of `mib` MiB."""
    gbs = ctypes.c_double()
    _check(cpp_lib.sr_measure_stream_read(int(mib) << 20, int(reps), ctypes.byref(gbs)), "measure_stream_read")
    return gbs.value


_sig("sr_measure_valu_issue", [_i, ctypes.POINTER(ctypes.c_double)])


def measure_valu_issue(reps=5):
    """Attainable vector-instruction issue rate (G wave-instructions/s), every SIMD full."""
    g = ctypes.c_double()
    _check(cpp_lib.sr_measure_valu_issue(int(reps), ctypes.byref(g)), "measure_valu_issue")
    return g.value


_sig("sr_measure_stream_write", [_l, _i, _i, ctypes.POINTER(ctypes.c_double)])


def measure_stream_write(mib=1024, reps=20, nontemporal=True):
    """Attainable HBM write bandwidth (GB/s): a plain streaming write of `mib` MiB."""
    gbs = ctypes.c_double()
    _check(cpp_lib.sr_measure_stream_write(int(mib) << 20, int(reps), 1 if nontemporal else 0, ctypes.byref(gbs)),
           "measure_stream_write")
    return gbs.value


_sig("sr_set_transfer_mode", [_i])
_sig("sr_transfer_mode", [])
KERNEL_FAMILIES.update({"build_s": 6, "gemm": 7, "build_b": 8, "gated_fix": 9, "near": 10})


_sig("sr_set_gather_mode", [_i])
_sig("sr_gather_mode", [])


def set_gather_mode(mode):
    """Plain fragment->fragment: 0 FP32 matrix-core products from 16 planes on, 1 vector units only."""
    _check(cpp_lib.sr_set_gather_mode(int(mode)), "set_gather_mode")


def set_transfer_mode(mode):
    """0: plain FP32 gathers; 1: FP16 MFMA transfer matrix (FP32 accumulation)."""
    _check(cpp_lib.sr_set_transfer_mode(int(mode)), "set_transfer_mode")


_sig("sr_set_transfer_splits", [_i])
_sig("sr_transfer_splits", [])


def set_transfer_splits(n):
    """Split-K of the transfer product (0: automatic); applies at the next preparation."""
    _check(cpp_lib.sr_set_transfer_splits(int(n)), "set_transfer_splits")


_sig("sr_set_transfer_symmetric", [_i])
_sig("sr_transfer_symmetric", [])


def set_transfer_symmetric(on):
    """2 (default): S from the blocks on and above the diagonal and their transposes
    where the layout allows, with the object's own visibility evaluated in that build
    when the scene has at most 4 occluder triangles; 1: the same, visibility always
    from the SELF visibility words; 0: every block evaluated. The same matrix every way."""
    _check(cpp_lib.sr_set_transfer_symmetric(int(on)), "set_transfer_symmetric")


_sig("sr_set_inline_visibility", [_i])
_sig("sr_inline_visibility", [])


def set_inline_visibility(on):
    """1 (default): with one incidence plane per source and at most 4 occluder
    triangles, the point->fragment sum evaluates its pairs' visibility itself;
    0: from the sources->fragments visibility words. The same result either way."""
    _check(cpp_lib.sr_set_inline_visibility(int(on)), "set_inline_visibility")


_sig("sr_set_reduce_in_product", [_i])
_sig("sr_reduce_in_product", [])


def set_reduce_in_product(on):
    """1: the transfer product's last workgroup of each row group reduces the split-K
    partial sums; 0 (default): a reduction pass of its own (faster on MI355X). The
    same result."""
    _check(cpp_lib.sr_set_reduce_in_product(int(on)), "set_reduce_in_product")


_sig("sr_read_stl", [ctypes.c_char_p, _fp, _l], _l)
_sig("sr_write_stl", [ctypes.c_char_p, _fp, _i])
_sig("sr_load_stl", [ctypes.c_char_p, _i, _f])
_sig("sr_colour_map", [_dp, _l, _f, _f, _i, _fp])
_sig("sr_surface_rgba", [_i, _i, _i, _fp, _l])


def read_stl(path):
    """(N, 3, 3) float32 triangles of a binary or ASCII STL file."""
    n = _check(cpp_lib.sr_read_stl(str(path).encode(), None, 0), "read_stl")
    out = np.zeros((max(n, 1), 9), np.float32)
    _check(cpp_lib.sr_read_stl(str(path).encode(), _ptr(out), n), "read_stl")
    return out[:n].reshape(-1, 3, 3)


def write_stl(path, tris):
    v = _f32(tris).reshape(-1, 9)
    _check(cpp_lib.sr_write_stl(str(path).encode(), _ptr(v), len(v)), "write_stl")


def load_stl(path, object_type, resolution):
    _check(cpp_lib.sr_load_stl(str(path).encode(), int(object_type), resolution), "load_stl")


def colour_map(p, min_db=-90.0, max_db=-20.0, phase=False):
    """RGBA (N, 4) of complex pressures, as the reference's surface render colours them."""
    p = np.ascontiguousarray(np.stack([np.real(p), np.imag(p)], -1).astype(np.float64).reshape(-1, 2))
    out = np.zeros((len(p), 4), np.float32)
    _check(cpp_lib.sr_colour_map(_ptr(p, _dp), len(p), min_db, max_db, 1 if phase else 0, _ptr(out)), "colour_map")
    return out


def surface_rgba(object_type, index, plane=0):
    n = cell_count(object_type, index)
    out = np.zeros((max(n, 1), 4), np.float32)
    _check(cpp_lib.sr_surface_rgba(object_type, index, plane, _ptr(out), n), "surface_rgba")
    return out[:n]


_sig("sr_set_near_mode", [_i, _f])
_sig("sr_nf_init", [ctypes.c_uint])
_sig("sr_nf_train", [_i, _i, _i, _f, ctypes.c_uint, _fp])
_sig("sr_nf_get_params", [_fp, _i])
_sig("sr_nf_set_params", [_fp, _i])
_sig("sr_nf_samples", [ctypes.c_uint, _i, _fp, _fp])
_sig("sr_nf_forward", [_fp, _i, _fp])
_sig("sr_nf_gradient", [_fp, _fp, _fp, _i, _fp])

NF_PARAMS = 4994
NF_IN = 10
NEAR_POINT, NEAR_QUADRATURE, NEAR_LEARNED = 0, 1, 2


def set_near_mode(mode, radius_wavelengths=2.0):
    _check(cpp_lib.sr_set_near_mode(int(mode), float(radius_wavelengths)), "set_near_mode")


_sig("sr_set_near_quadrature_radius", [_f])


def set_near_quadrature_radius(deltas):
    """Learned mode: pairs closer than `deltas` fragment sizes (default 3, the network's
    training range starts there) use the 16x16 quadrature."""
    _check(cpp_lib.sr_set_near_quadrature_radius(float(deltas)), "set_near_quadrature_radius")


def nf_init(seed=1):
    _check(cpp_lib.sr_nf_init(seed), "nf_init")


def nf_train(nsamples=262144, steps=3000, batch=4096, lr=3e-3, seed=7):
    rep = np.zeros(2, np.float32)
    _check(cpp_lib.sr_nf_train(nsamples, steps, batch, lr, seed, _ptr(rep)), "nf_train")
    return float(rep[0]), float(rep[1])


def nf_params():
    out = np.zeros(NF_PARAMS, np.float32)
    _check(cpp_lib.sr_nf_get_params(_ptr(out), NF_PARAMS), "nf_get_params")
    return out


def nf_set_params(p):
    p = _f32(p).ravel()
    _check(cpp_lib.sr_nf_set_params(_ptr(p), len(p)), "nf_set_params")


def nf_samples(seed, n):
    X = np.zeros((n, NF_IN), np.float32)
    Y = np.zeros((n, 2), np.float32)
    _check(cpp_lib.sr_nf_samples(seed, n, _ptr(X), _ptr(Y)), "nf_samples")
    return X, Y


def nf_forward(X):
    X = _f32(X).reshape(-1, NF_IN)
    out = np.zeros((len(X), 2), np.float32)
    _check(cpp_lib.sr_nf_forward(_ptr(X), len(X), _ptr(out)), "nf_forward")
    return out


def nf_gradient(params, X, Y):
    p, X, Y = _f32(params).ravel(), _f32(X).reshape(-1, NF_IN), _f32(Y).reshape(-1, 2)
    g = np.zeros(NF_PARAMS, np.float32)
    _check(cpp_lib.sr_nf_gradient(_ptr(p), _ptr(X), _ptr(Y), len(X), _ptr(g)), "nf_gradient")
    return g


_sig("sr_device_synchronize", [])


def device_synchronize():
    """Waits for all work of the library's HIP runtime on the current device."""
    _check(cpp_lib.sr_device_synchronize(), "device_synchronize")


_sig("sr_context_create", [])
_sig("sr_context_select", [_i])
_sig("sr_context_current", [])
_sig("sr_context_destroy", [_i])


def context_create():
    """A new, empty scene context; returns its handle (> 0). Every other call acts on
    the calling thread's current context (handle 0, the default, until one is selected)."""
    h = cpp_lib.sr_context_create()
    _check(h, "context_create")
    return h


def context_select(handle):
    """Makes handle current for the calling thread (0: the default context)."""
    _check(cpp_lib.sr_context_select(int(handle)), "context_select")


def context_current():
    return cpp_lib.sr_context_current()


def context_destroy(handle):
    """Tears a created context down and frees it (not 0, not the current one)."""
    _check(cpp_lib.sr_context_destroy(int(handle)), "context_destroy")


for _n, _a in [("sr_batch_clear", []), ("sr_batch_add_scene", [_fp, _i, _f, _f, _f, _f, _fp, _i, _fp, _i]),
               ("sr_batch_set_bounces", [_i]), ("sr_batch_prepare", []),
               ("sr_batch_run", [_i, _fp]), ("sr_batch_autotune", [_fp, _i]),
               ("sr_batch_field_points", [_i, _dp, _i]), ("sr_batch_get_tune", [_ip]), ("sr_batch_set_tune", [_ip])]:
    _sig(_n, _a)
_sig("sr_batch_pairs", [], ctypes.c_double)


def batch_clear():
    _check(cpp_lib.sr_batch_clear(), "batch_clear")


def batch_add_scene(tris, delta, f, sources, points, c=1480.0, atten=0.0):
    t = _f32(tris).reshape(-1, 9)
    s = _f32(sources).reshape(-1, 3)
    p = _f32(points).reshape(-1, 3)
    return _check(cpp_lib.sr_batch_add_scene(_ptr(t), len(t), delta, c, f, atten, _ptr(s), len(s), _ptr(p), len(p)),
                  "batch_add_scene")


def batch_set_bounces(n):
    _check(cpp_lib.sr_batch_set_bounces(int(n)), "batch_set_bounces")


def batch_set_near(mode, radius_wavelengths=2.0, quad_deltas=3.0):
    """Near field of the batched bounces: 0 point rule, 1 16x16 quadrature, 2 learned."""
    _check(cpp_lib.sr_batch_set_near(int(mode), float(radius_wavelengths), float(quad_deltas)), "batch_set_near")


def batch_prepare():
    _check(cpp_lib.sr_batch_prepare(), "batch_prepare")


def batch_run(use_graph=True):
    ms = np.zeros(1, np.float32)
    _check(cpp_lib.sr_batch_run(1 if use_graph else 0, _ptr(ms)), "batch_run")
    return float(ms[0])


def batch_autotune(nmax=64):
    rep = np.zeros((nmax, 5), np.float32)
    k = _check(cpp_lib.sr_batch_autotune(_ptr(rep), nmax), "batch_autotune")
    return rep[:k]


def batch_field_points(scene, n):
    out = np.zeros((n, 2), np.float64)
    _check(cpp_lib.sr_batch_field_points(scene, _ptr(out, _dp), n), "batch_field_points")
    return out[:, 0] + 1j * out[:, 1]


def batch_pairs():
    return float(cpp_lib.sr_batch_pairs())


def batch_tune(values=None):
    v = np.zeros(4, np.int32)
    if values is not None:
        v[:] = values
        _check(cpp_lib.sr_batch_set_tune(_ptr(v, _ip)), "batch_set_tune")
    _check(cpp_lib.sr_batch_get_tune(_ptr(v, _ip)), "batch_get_tune")
    return v.tolist()
