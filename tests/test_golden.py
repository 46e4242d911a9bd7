"""The oracle against its committed fixtures (tests/golden/c1_golden.json,
written by tests/golden/make_golden.py)."""
import hashlib
import json
import os

import numpy as np

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
