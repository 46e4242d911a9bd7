"""The oracle against its committed fixtures (tests/golden/c1_golden.json,
written by tests/golden/make_golden.py; tests/golden/c1_smoothers.json, written
by tests/golden/make_c1_smoothers.py), and the library's host topologist
against the restructure fixture."""
import hashlib
import json
import os

import numpy as np
import pytest

from gpumesh_amd import cube_mesh

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "c1_golden.json")))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c1_input_and_topology(oracle):
    m = cube_mesh(20, "tet", 0.3, seed=1)
    assert digest(m.pos) == GOLDEN["input_positions_sha256"]
    om = oracle.OracleMesh.from_cube(m)
    assert digest(om.permutation()) == GOLDEN["permutation_sha256"]
    groups, _ = om.groups()
    assert groups[:, 4:6].tolist() == GOLDEN["group_interior_ranges"]


def test_c1_nelder_mead_both_libm_modes(oracle):
    m = cube_mesh(20, "tet", 0.3, seed=1)
    for mode, key in ((0, "glibc"), (1, "gm_libm")):
        oracle.set_libm(mode)
        try:
            om = oracle.OracleMesh.from_cube(m)
            r = om.smooth("nm", 10, threads=8)
        finally:
            oracle.set_libm(1)
        g = GOLDEN[key]
        assert r["min"].tolist() == g["min"]
        assert np.allclose(r["mean"], g["mean"], rtol=0, atol=1e-15)
        assert int(r["eval_counts"].sum()) == g["eval_count_total"]
        assert digest(om.positions()) == g["final_positions_sha256"]


SMOOTHERS = json.load(open(os.path.join(HERE, "golden", "c1_smoothers.json")))


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_c1_every_smoother_against_fixture(oracle, smoother):
    """Per-pass records and the positions after passes 1, 2 and 10
    (tests/golden/make_c1_smoothers.py)."""
    g = SMOOTHERS["smoothers"][smoother]
    om = oracle.OracleMesh.from_cube(cube_mesh(20, "tet", 0.3, seed=1))
    mins, evals = [], 0
    for upto, step in ((1, 1), (2, 1), (10, 8)):
        r = om.smooth(smoother, step, threads=8)
        mins += r["min"].tolist() if not mins else r["min"][1:].tolist()
        evals += int(r["eval_counts"].sum())
        assert digest(om.positions()) == g["positions_sha256"][str(upto)], upto
        assert r["hist"][-1].tolist() == g["hist"][str(upto)], upto
    assert mins == g["min"] and evals == g["eval_count_total"]


def test_c1_restructure_against_fixture(oracle):
    """The oracle's BATR and the library's host topologist on the C1 mesh."""
    from gpumesh_amd import GpuMesh
    g = SMOOTHERS["restructure"]
    m = cube_mesh(20, "tet", 0.3, seed=1)
    om = oracle.OracleMesh.from_cube(m)
    gm = GpuMesh.from_cube(m, device=False)
    for mesh in (om, gm):
        mesh.set_crew("mean_ratio", "uniform", scaling=g["scaling"])
        counts = mesh.restructure(g["topo_passes"], g["refinement_sweeps"], g["min_edge"], g["max_edge"])
        assert counts == g["counts"]
        assert digest(mesh.positions()) == g["positions_sha256"]
        assert digest(mesh.elements()[0]) == g["tets_sha256"]
