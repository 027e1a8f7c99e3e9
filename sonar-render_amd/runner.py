"""Runs a scene dict (see scenes.py) through the GPU engine's C-ABI.

The call sequence is the reference's production one
(run_scripts/modules/mh.py:43-70 plus ProjectTargetToFieldPoints):
load points -> set_initial_conditions -> load meshes -> render_cuda ->
project sources -> bounces -> targets to field objects -> targets to field points.
"""
import time

import numpy as np

import sonar_api as api


def load_scene(scene, batched=None, transfer_mode=0):
    api.TearDownCuda()
    if batched is None:
        batched = scene.get("batched", False)
    api.set_batched_sources(batched)
    api.load_points_to_cuda(scene["sources"], isSource=True)
    api.load_points_to_cuda(scene["points"], isSource=False)
    api.set_initial_conditions(scene["c"], scene["f"], scene.get("atten", 0.0))
    for tris, d in scene["targets"]:
        api.load_stl_mesh_to_cuda(tris, 0, d)
    for tris, d in scene.get("fields", []):
        api.load_stl_mesh_to_cuda(tris, 2, d)
    for tris, d in scene.get("source_objects", []):
        api.load_stl_mesh_to_cuda(tris, 1, d)
    api.render_cuda()
    api.set_transfer_mode(transfer_mode)


def project_scene(scene, timings=None):
    steps = [("src", lambda: api.project_source_points_to_objects(scene.get("project_field_objects", False)))]
    if scene.get("bounces", 0):
        steps.append(("bounces", lambda: api.project_target_to_target_objects(scene["bounces"])))
    if scene.get("fields"):
        steps.append(("field_objects", api.project_target_to_field_objects))
    steps.append(("field_points", api.project_target_to_field_points))
    for name, fn in steps:
        t0 = time.perf_counter()
        fn()
        if timings is not None:
            timings[name] = timings.get(name, 0.0) + (time.perf_counter() - t0)


def read_results(scene):
    nq = len(scene["points"])
    planes = api.num_planes()
    return dict(
        points=np.stack([api.field_points(nq, p) for p in range(planes)]),
        targets=[np.stack([api.object_pressures(0, i, 0, p) for p in range(planes)])
                 for i in range(len(scene["targets"]))],
        fields=[np.stack([api.object_pressures(2, i, 0, p) for p in range(planes)])
                for i in range(len(scene.get("fields", [])))],
    )


def run_scene(scene, batched=None, keep=False, transfer_mode=0):
    load_scene(scene, batched, transfer_mode)
    timings = {}
    project_scene(scene, timings)
    out = read_results(scene)
    out["timings"] = timings
    if not keep:
        api.TearDownCuda()
    return out


def facet_active_counts(tris, delta):
    """Active (A > 1e-6) cell count of each facet, from the library's host fragmenter."""
    counts, cells = [], 0
    for t in np.asarray(tris, np.float32).reshape(-1, 3, 3):
        (nx, ny, _), area, _, _ = api.fragment_facet(t, delta)
        counts.append(int((area > 1e-6).sum()))
        cells += nx * ny
    return counts, cells


def pair_count(scene):
    """Ordered pairs of one scene run (see scenes.scene_pairs)."""
    import scenes
    tg = []
    for tris, d in scene["targets"]:
        per, cells = facet_active_counts(tris, d)
        tg.append((sum(per), cells, per))
    fl = [sum(facet_active_counts(tris, d)[0]) for tris, d in scene.get("fields", [])]
    return scenes.scene_pairs(scene, tg, fl)
