"""ctypes wrapper of the FP64 host oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle.so")


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


if not os.path.exists(LIB):
    build()
_lib = ctypes.CDLL(LIB)

_v = ctypes.c_void_p
_f = ctypes.c_float
_i = ctypes.c_int
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_u8p = ctypes.POINTER(ctypes.c_ubyte)

for name, args, res in [
    ("orc_create", [], _v), ("orc_destroy", [_v], None), ("orc_set_threads", [_i], None),
    ("orc_get_threads", [], _i), ("orc_set_bruteforce", [_v, _i], None),
    ("orc_set_initial_conditions", [_v, _f, _f, _f, _f], None),
    ("orc_load_geometry", [_v, _fp, _i, _i, _f], _i),
    ("orc_load_source_points", [_v, _fp, _i], None), ("orc_load_field_points", [_v, _fp, _i], None),
    ("orc_set_source_amplitude", [_v, _i, ctypes.c_double, ctypes.c_double], None),
    ("orc_render", [_v], _i), ("orc_project_src_to_objects", [_v, _i], _i),
    ("orc_project_target_to_target", [_v, _i], _i), ("orc_project_resp", [_v, _i, _i], _i),
    ("orc_target_pout_to_pin", [_v, _i], _i),
    ("orc_project_target_to_field_objects", [_v], _i), ("orc_project_target_to_field_points", [_v], _i),
    ("orc_num_field_points", [_v], _i), ("orc_get_field_points", [_v, _dp], None),
    ("orc_object_ncell", [_v, _i, _i], ctypes.c_long), ("orc_object_nfacet", [_v, _i, _i], _i),
    ("orc_object_grids", [_v, _i, _i, _ip], None),
    ("orc_object_cells", [_v, _i, _i, _fp, _fp, _fp, _ip], None),
    ("orc_object_frames", [_v, _i, _i, _fp], None),
    ("orc_get_object_pressure", [_v, _i, _i, _i, _dp], None),
    ("orc_set_object_pressure", [_v, _i, _i, _i, _dp], None),
    ("orc_anyhit", [_v, _fp, _fp, _fp, _fp, ctypes.c_long, _u8p, _i, _ip, _ip], None),
    ("orc_vis_params", [_v, _fp], None), ("orc_kf", [_v], _f),
    ("orc_set_near_quadrature", [_v, _i, _f], None),
    ("orc_set_source_object_amplitude", [_v, _i, ctypes.c_double, ctypes.c_double], None),
    ("orc_near_pair", [ctypes.c_double, ctypes.c_double, _i, _f, _f, _fp, _fp, _fp, _fp, _fp, _fp, _dp], None),
    ("orc_cosine_kept", [_fp, _fp, ctypes.c_long, _u8p], None),
]:
    fn = getattr(_lib, name)
    fn.argtypes = args
    fn.restype = res


def near_pair(kr, atten, nsub, delta, area, pe, xa, ya, na, nb, pr):
    """One pair of the near-field judge (nsub x nsub sub-fragment quadrature of the
    pair kernel over the emitter's delta x delta cell); nb None: field destination."""
    out = np.zeros(2)
    a = [_f32(q).reshape(3) for q in (pe, xa, ya, na)]
    nbp = None if nb is None else _f32(nb).reshape(3)
    prp = _f32(pr).reshape(3)
    _lib.orc_near_pair(float(kr), float(atten), int(nsub), float(delta), float(area),
                       *[q.ctypes.data_as(_fp) for q in a],
                       None if nbp is None else nbp.ctypes.data_as(_fp), prp.ctypes.data_as(_fp),
                       out.ctypes.data_as(_dp))
    return complex(out[0], out[1])


def cosine_kept(v, n):
    """The oracle's cosine cull of pairs (difference vectors v, unit normals n, FP32):
    True where the pair is kept (cos formed by ray_dir and dotp is not below 1e-6)."""
    v, n = _f32(v).reshape(-1, 3), _f32(n).reshape(-1, 3)
    out = np.zeros(len(v), np.uint8)
    _lib.orc_cosine_kept(v.ctypes.data_as(_fp), n.ctypes.data_as(_fp), len(v), out.ctypes.data_as(_u8p))
    return out.astype(bool)


def set_threads(n):
    _lib.orc_set_threads(int(n))


def get_threads():
    return _lib.orc_get_threads()


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class Scene:
    """One oracle scene; methods mirror the reference's call sequence."""

    def __init__(self):
        self.h = _lib.orc_create()

    def close(self):
        if self.h:
            _lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_initial_conditions(self, cp, f, att=0.0, density=0.0):
        _lib.orc_set_initial_conditions(self.h, cp, f, att, density)

    def load_geometry(self, tris, obj_type, res):
        v = _f32(tris).reshape(-1, 9)
        return _lib.orc_load_geometry(self.h, v.ctypes.data_as(_fp), len(v), obj_type, res)

    def load_source_points(self, p):
        p = _f32(p).reshape(-1, 3)
        _lib.orc_load_source_points(self.h, p.ctypes.data_as(_fp), len(p))

    def load_field_points(self, p):
        p = _f32(p).reshape(-1, 3)
        _lib.orc_load_field_points(self.h, p.ctypes.data_as(_fp), len(p))

    def set_source_amplitude(self, i, v):
        _lib.orc_set_source_amplitude(self.h, i, float(np.real(v)), float(np.imag(v)))

    def set_source_object_amplitude(self, i, v):
        _lib.orc_set_source_object_amplitude(self.h, i, float(np.real(v)), float(np.imag(v)))

    def set_bruteforce(self, on):
        _lib.orc_set_bruteforce(self.h, 1 if on else 0)

    def set_near_quadrature(self, nsub, radius):
        """Near-field judge: pairs closer than radius use nsub x nsub sub-fragment quadrature."""
        _lib.orc_set_near_quadrature(self.h, int(nsub), float(radius))

    def render(self):
        assert _lib.orc_render(self.h) == 0

    def project_src(self, field_objects=False):
        assert _lib.orc_project_src_to_objects(self.h, 1 if field_objects else 0) == 0

    def project_target_to_target(self, n):
        assert _lib.orc_project_target_to_target(self.h, n) == 0

    def project_target_to_field_objects(self):
        assert _lib.orc_project_target_to_field_objects(self.h) == 0

    def project_target_to_field_points(self):
        assert _lib.orc_project_target_to_field_points(self.h) == 0

    def prepare_resp(self, t):
        assert _lib.orc_target_pout_to_pin(self.h, t) == 0

    def project_resp(self, a, b):
        assert _lib.orc_project_resp(self.h, a, b) == 0

    def field_points(self):
        n = _lib.orc_num_field_points(self.h)
        out = np.zeros((n, 2), np.float64)
        _lib.orc_get_field_points(self.h, out.ctypes.data_as(_dp))
        return out[:, 0] + 1j * out[:, 1]

    def ncell(self, t, i):
        return _lib.orc_object_ncell(self.h, t, i)

    def grids(self, t, i):
        nf = _lib.orc_object_nfacet(self.h, t, i)
        g = np.zeros((nf, 3), np.int32)
        _lib.orc_object_grids(self.h, t, i, g.ctypes.data_as(_ip))
        return g

    def cells(self, t, i):
        n = self.ncell(t, i)
        area = np.zeros(n, np.float32)
        pe = np.zeros((n, 3), np.float32)
        pr = np.zeros((n, 3), np.float32)
        fid = np.zeros(n, np.int32)
        _lib.orc_object_cells(self.h, t, i, area.ctypes.data_as(_fp), pe.ctypes.data_as(_fp),
                              pr.ctypes.data_as(_fp), fid.ctypes.data_as(_ip))
        return dict(area=area, pe=pe, pr=pr, facet=fid)

    def frames(self, t, i):
        nf = _lib.orc_object_nfacet(self.h, t, i)
        out = np.zeros((nf, 16), np.float32)
        _lib.orc_object_frames(self.h, t, i, out.ctypes.data_as(_fp))
        return out

    def pressure(self, t, i, which=0):
        n = self.ncell(t, i)
        out = np.zeros((n, 2), np.float64)
        _lib.orc_get_object_pressure(self.h, t, i, which, out.ctypes.data_as(_dp))
        return out[:, 0] + 1j * out[:, 1]

    def set_pressure(self, t, i, which, values):
        v = np.ascontiguousarray(np.stack([np.real(values), np.imag(values)], 1).astype(np.float64))
        _lib.orc_set_object_pressure(self.h, t, i, which, v.ctypes.data_as(_dp))

    def anyhit(self, o, d, tmin, tmax, brute=False, skip_a=None, skip_b=None):
        """skip_a, skip_b: per-ray occluder triangles the end points lie on (never tested)."""
        o, d = _f32(o).reshape(-1, 3), _f32(d).reshape(-1, 3)
        n = len(o)
        tmin = _f32(np.broadcast_to(tmin, (n,)))
        tmax = _f32(np.broadcast_to(tmax, (n,)))
        out = np.zeros(n, np.uint8)
        sa = None if skip_a is None else np.ascontiguousarray(np.broadcast_to(skip_a, (n,)), np.int32)
        sb = None if skip_b is None else np.ascontiguousarray(np.broadcast_to(skip_b, (n,)), np.int32)
        _lib.orc_anyhit(self.h, o.ctypes.data_as(_fp), d.ctypes.data_as(_fp), tmin.ctypes.data_as(_fp),
                        tmax.ctypes.data_as(_fp), n, out.ctypes.data_as(_u8p), 1 if brute else 0,
                        None if sa is None else sa.ctypes.data_as(_ip), None if sb is None else sb.ctypes.data_as(_ip))
        return out.astype(bool)

    def vis_params(self):
        out = np.zeros(5, np.float32)
        _lib.orc_vis_params(self.h, out.ctypes.data_as(_fp))
        return out

    def kf(self):
        return _lib.orc_kf(self.h)


def run_scene(scene, plane_source=None, bruteforce=False, near=None):
    """Runs a scene dict (see scenes.py) through the oracle.

    plane_source: for batched scenes, the single source index to run.
    Returns dict(points=complex128[Q], targets=[P per cell], fields=[P per cell]).
    """
    s = Scene()
    s.set_bruteforce(bruteforce)
    if near:
        s.set_near_quadrature(*near)
    s.set_initial_conditions(scene["c"], scene["f"], scene.get("atten", 0.0))
    for tris, d in scene["targets"]:
        s.load_geometry(tris, 0, d)
    for tris, d in scene.get("fields", []):
        s.load_geometry(tris, 2, d)
    for tris, d in scene.get("source_objects", []):
        s.load_geometry(tris, 1, d)
    src = scene["sources"] if plane_source is None else scene["sources"][plane_source:plane_source + 1]
    s.load_source_points(src)
    s.load_field_points(scene["points"])
    s.render()
    s.project_src(scene.get("project_field_objects", False))
    if scene.get("bounces", 0):
        s.project_target_to_target(scene["bounces"])
    if scene.get("fields"):
        s.project_target_to_field_objects()
    s.project_target_to_field_points()
    out = dict(points=s.field_points(),
               targets=[s.pressure(0, i) for i in range(len(scene["targets"]))],
               fields=[s.pressure(2, i) for i in range(len(scene.get("fields", [])))])
    s.close()
    return out
