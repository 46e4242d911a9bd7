"""Degenerate inputs on the GPU: meshes without a movable vertex, zero
passes, and such meshes inside a batch.

A one-cell cube has every vertex on the boundary, so it has no colour group:
the passes launch only the element steps (GETMe) and the evaluation, the
positions must not move, and every record equals the oracle's.  Zero passes
give the one record of the starting mesh.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuBatch, GpuMesh, cube_mesh

pytestmark = pytest.mark.gpu

COEFF = {"nm": 1.0 / 20.0, "gd": 1.0 / 20.0, "spawn": 0.25, "getme": 1.0 / 20.0}


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_polish"])
def test_no_movable_vertex(oracle, smoother, kind, precision):
    m = cube_mesh(1, kind, 0.0, seed=1)
    assert m.fixed.all()
    om = oracle.OracleMesh.from_cube(m)
    gm = GpuMesh.from_cube(m)
    assert gm.ngroups == 0 and gm.nmovable == 0
    before = gm.positions()
    ref = om.smooth(smoother, 2, threads=1, spawn_coeff=COEFF[smoother])
    out = gm.smooth(smoother, 2, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother])
    assert np.array_equal(gm.positions(), before)
    assert not out["eval_counts"].any()
    if precision == "fp64":
        assert np.array_equal(out["min"], ref["min"])
        assert np.array_equal(out["hist"], ref["hist"])
    else:
        assert np.max(np.abs(out["min"] - ref["min"])) < 1e-6


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_zero_passes_is_the_start_record(oracle, smoother):
    m = cube_mesh(4, "tet", 0.3, seed=2)
    om = oracle.OracleMesh.from_cube(m)
    gm = GpuMesh.from_cube(m)
    before = gm.positions()
    ref = om.smooth(smoother, 0, threads=1, spawn_coeff=COEFF[smoother])
    out = gm.smooth(smoother, 0, precision="fp64", kernel="wave", spawn_coeff=COEFF[smoother])
    assert out["min"].shape == (1,) and np.array_equal(out["min"], ref["min"])
    assert np.array_equal(out["hist"], ref["hist"])
    assert np.array_equal(gm.positions(), before)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32_polish"])
@pytest.mark.parametrize("smoother", ["gd", "getme"])
def test_batch_with_meshes_without_movable_vertices(smoother, precision, graph):
    """Members without a colour group next to ordinary ones: every member
    comes out bit-identical to the same mesh smoothed alone."""
    specs = [(1, "tet", 1), (5, "tet", 3), (1, "hex", 1), (4, "hex", 4)]
    meshes = [cube_mesh(n, kind, 0.3 if n > 1 else 0.0, seed=s) for n, kind, s in specs]
    alone = []
    for m in meshes:
        g = GpuMesh.from_cube(m)
        r = g.smooth(smoother, 3, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother])
        alone.append((g.positions(), r))
    members = [GpuMesh.from_cube(m) for m in meshes]
    b = GpuBatch(members)
    rb = b.smooth(smoother, 3, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother], graph=graph)
    for i, (pos, r) in enumerate(alone):
        assert np.array_equal(members[i].positions().view(np.uint64), pos.view(np.uint64)), i
        assert np.array_equal(rb["min"][i].view(np.uint64), r["min"].view(np.uint64)), i
        assert np.array_equal(rb["hist"][i], r["hist"]), i


def test_metric_wise_zero_length_edge(oracle):
    """Two coincident vertices under the metric-wise measurer: the adaptive
    walk ends on the NaN pair difference; the elements holding the edge get
    NaN, the others their qualities, bit for bit with the oracle."""
    m = cube_mesh(2, "tet", 0.0, seed=1)
    inner = int(np.where(m.fixed == 0)[0][0])
    t = m.tets[np.any(m.tets == inner, axis=1)][0]
    pos = m.pos.copy()
    pos[inner] = pos[int(t[t != inner][0])]
    om = oracle.OracleMesh(pos, m.fixed, m.tets, m.hexs)
    gm = GpuMesh(pos, m.fixed, m.tets, m.hexs)
    om.set_crew("mean_ratio", "analytic", scaling=2.0)
    gm.set_crew("mean_ratio", "analytic", scaling=2.0)
    g = gm.evaluate(element_qualities=True)
    q = om.element_qualities()
    assert np.isnan(q).sum() == 6
    assert np.array_equal(g["elem_qual"], q, equal_nan=True)
    r = om.evaluate()
    assert g["min"] == r["min"]
    assert np.array_equal(g["hist"], r["hist"])


def test_throughput_probe_is_sane():
    """gm_throughput_probe: positive rates, transcendentals no faster than
    FMAs, and the mixed loop no faster than its slower part."""
    from gpumesh_amd._lib import check, lib
    r = [np.zeros(3) for _ in range(3)]
    for op in range(3):
        check(lib().gm_throughput_probe(op, r[op].ctypes.data), "gm_throughput_probe")
    trans, fma = r[0][0], r[1][1]
    assert 1e12 < trans <= 1.1 * fma < 1e15
    assert r[0][1] == 0.0 and r[1][0] == 0.0
    assert r[2][2] >= 0.9 * max(r[0][2], r[1][2])
