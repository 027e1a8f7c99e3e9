"""GPU: the plate-thick trihedral (tests/golden/thick_trihedral.stl, the geometry of the reference's
run_scripts/trihedral_reflector.stl rebuilt by tests/golden/make_thick_trihedral.py: 36 triangles of a
closed plate-thick corner reflector, so back faces and shadowing matter) through the plain form (vector
units and matrix cores) and the transfer form, against the FP64 oracle."""
import os

import numpy as np
import pytest

import oracle
import runner
import scenes
from conftest import rel_l2

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "thick_trihedral.stl")


def _scene(api, n_side):
    tris = api.read_stl(FIXTURE)
    f = 3e3
    src = scenes.octant_incidences(n_side, 12.0, seed=4)
    return dict(name="thick_trihedral_mesh", f=f, c=1480.0, atten=0.0, targets=[(tris, 1480.0 / f / 7.5)],
                fields=[], sources=src, points=src.copy(), bounces=2, project_field_objects=False, batched=True)


@pytest.mark.parametrize("transfer_mode,gather_mode,n_side", [(0, 1, 2), (0, 0, 4), (1, 0, 2)])
def test_reference_mesh_vs_oracle(api, transfer_mode, gather_mode, n_side):
    sc = _scene(api, n_side)
    print(f"[gpu] thick trihedral mesh: transfer {transfer_mode}, gather {gather_mode}, "
          f"{len(sc['sources'])} incidences", flush=True)
    runner.load_scene(sc, transfer_mode=transfer_mode)
    api.set_gather_mode(gather_mode)
    runner.project_scene(sc)
    out = runner.read_results(sc)
    api.TearDownCuda()
    tol = 2e-4 if transfer_mode == 0 else 1e-3
    for p in (0, len(sc["sources"]) - 1):
        ref = oracle.run_scene(sc, plane_source=p)
        e = rel_l2(np.concatenate([out["points"][p], out["targets"][0][p]]),
                   np.concatenate([ref["points"], ref["targets"][0]]))
        print("plane", p, "rel-L2", e)
        assert e < tol
