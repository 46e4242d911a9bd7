"""Sliding boundary vertices on the GPU (gm_mesh_create_desc): BoxBoundary
meshes on [-1, 1]^3 (Boundaries/BoxBoundary.cpp:7-36) whose corners are
fixed, edge vertices slide on the box's 12 edges and face vertices on its 6
faces, smoothed by the FP64 kernels and compared with the oracle's Thread
sweep (AbstractVertexWiseSmoother.cpp:84-152, snapping through
MeshTopo::snapToBoundary, DataStructures/Mesh.h:256-267): positions, patch
evaluation counts, minimum and histogram bit-identical, harmonic mean within
1e-12 (a different summation order)."""
import numpy as np
import pytest

from gpumesh_amd import GpuBatch, GpuMesh, box_boundary, cube_mesh

pytestmark = pytest.mark.gpu

MEAN_TOL = 1e-12


def _box(kind, n, seed=5, jitter=0.25):
    m = cube_mesh(n, kind, jitter, seed=seed)
    pos, fixed, cons, constraints = box_boundary(m, 0.2, seed)
    return m, pos, fixed, cons, constraints


def _pair(oracle, kind, n, seed=5):
    m, pos, fixed, cons, constraints = _box(kind, n, seed)
    om = oracle.OracleMesh(pos, fixed, m.tets, m.hexs, constraints=constraints, constraint=cons)
    gm = GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, constraints)
    return om, gm, (pos, fixed, cons, constraints)


def _same(out, ref, om, gm):
    p_ref, p_gpu = om.positions(), gm.positions()
    assert np.array_equal(p_gpu, p_ref), f"max |d| = {np.max(np.abs(p_gpu - p_ref))}"
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])
    assert np.max(np.abs(out["mean"] - ref["mean"])) <= MEAN_TOL
    assert np.array_equal(out["hist"], ref["hist"])


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kernel", ["plain", "wave"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_box_smoothing_fp64_bitexact(oracle, smoother, kernel, kind):
    n = 8 if kind == "tet" else 7
    om, gm, _ = _pair(oracle, kind, n)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    ref = om.smooth(smoother, 3, threads=4, spawn_coeff=coeff)
    out = gm.smooth(smoother, 3, precision="fp64", kernel=kernel, spawn_coeff=coeff)
    _same(out, ref, om, gm)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_larger_box_nm_fp64_bitexact(oracle, kind):
    """Many blocks per colour step, several constrained runs per colour."""
    om, gm, _ = _pair(oracle, kind, 16 if kind == "tet" else 13, seed=9)
    ref = om.smooth("nm", 2, threads=4)
    out = gm.smooth("nm", 2, precision="fp64", kernel="wave")
    _same(out, ref, om, gm)


@pytest.mark.parametrize("smoother", ["nm", "getme"])
def test_box_metric_conformity_fp64_bitexact(oracle, smoother):
    """The anisotropic crew (metric conformity, the reference's rotated
    analytic field) on a sliding hex box."""
    om, gm, _ = _pair(oracle, "hex", 5)
    for x in (om, gm):
        x.set_crew("metric_conformity", "analytic_reference", scaling=4.0)
        x.set_aspect_ratio(4.0)
    ref = om.smooth(smoother, 2, threads=4)
    out = gm.smooth(smoother, 2, precision="fp64", kernel="wave")
    _same(out, ref, om, gm)


def test_box_autopilot_fp64_bitexact(oracle):
    om, gm, _ = _pair(oracle, "tet", 7)
    ref = om.smooth_autopilot("nm", 6, min_qual_threshold=0.001, threads=4)
    out = gm.smooth_autopilot("nm", 6, min_qual_threshold=0.001, precision="fp64", kernel="wave")
    assert out["passes"] == len(ref["min"]) - 1
    assert np.array_equal(om.positions(), gm.positions())
    assert np.array_equal(out["min"], ref["min"])


def _on_box(p, pos, fixed, cons, constraints, tol):
    """Corners where they were; face and edge vertices on their faces and
    edges within tol; everything inside the box."""
    assert np.array_equal(p[fixed > 0], pos[fixed > 0])
    for v in np.where(cons >= 0)[0]:
        kind, a, b = constraints[cons[v]]
        if kind == 1:
            axis = int(np.argmax(np.abs(b)))
            assert abs(p[v, axis] - a[axis]) <= tol
        else:
            for i in range(3):
                if a[i] == b[i]:
                    assert abs(p[v, i] - a[i]) <= tol
    assert np.all(np.abs(p) <= 1.0 + tol)


@pytest.mark.parametrize("precision,tol", [("fp64_fast", 0.0), ("fp32_polish", 0.0), ("fp32", 1e-6)])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn"])
def test_box_mixed_precision_stays_on_the_box(oracle, smoother, precision, tol):
    """The faster precisions keep sliding vertices on their constraints and
    reach the FP64 sweep's quality (equal to 1e-6 in the harmonic mean)."""
    om, gm, (pos, fixed, cons, constraints) = _pair(oracle, "tet", 10)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    ref = om.smooth(smoother, 3, threads=4, spawn_coeff=coeff)
    out = gm.smooth(smoother, 3, precision=precision, kernel="wave", spawn_coeff=coeff)
    p = gm.positions()[gm.permutation()]
    _on_box(p, pos, fixed, cons, constraints, tol)
    assert abs(out["mean"][-1] - ref["mean"][-1]) <= 1e-4, (out["mean"][-1], ref["mean"][-1])
    assert out["mean"][-1] > out["mean"][0]


def test_box_batch_equals_solo(oracle):
    """Constrained members in one fused batch: each bit-identical to its
    own solo run, which equals the oracle."""
    members, solos, refs = [], [], []
    for i, (kind, n) in enumerate([("tet", 6), ("hex", 5), ("tet", 4)]):
        m, pos, fixed, cons, constraints = _box(kind, n, seed=20 + i)
        members.append(GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, constraints))
        solos.append(GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, constraints))
        refs.append(oracle.OracleMesh(pos, fixed, m.tets, m.hexs, constraints=constraints, constraint=cons))
    b = GpuBatch(members)
    out = b.smooth("nm", 2, precision="fp64", kernel="wave")
    for k in range(len(members)):
        solo = solos[k].smooth("nm", 2, precision="fp64", kernel="wave")
        ref = refs[k].smooth("nm", 2, threads=4)
        assert np.array_equal(members[k].positions(), solos[k].positions())
        assert np.array_equal(members[k].positions(), refs[k].positions())
        assert np.array_equal(out["min"][k], solo["min"]) and np.array_equal(out["min"][k], ref["min"])
    b.close()


def test_grouped_box_matches_oracle(oracle):
    """The caller's numbering and colour groups with sliding vertices: the
    oracle's own groups handed back in the input numbering."""
    m, pos, fixed, cons, constraints = _box("tet", 6)
    om = oracle.OracleMesh(pos, fixed, m.tets, m.hexs, constraints=constraints, constraint=cons)
    perm = om.permutation()                        # old -> new
    inv = np.argsort(perm)                         # new -> old
    groups, _ = om.groups()
    off, nodes = [0], []
    for g in range(len(groups)):
        ids = np.concatenate([np.arange(groups[g, 2 * t], groups[g, 2 * t + 1]) for t in range(3)])
        nodes.extend(inv[ids].tolist())
        off.append(len(nodes))
    gm = GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, constraints,
                             groups=(np.array(off), np.array(nodes)))
    ref = om.smooth("nm", 2, threads=4)
    out = gm.smooth("nm", 2, precision="fp64", kernel="wave")
    assert np.array_equal(gm.positions(), om.positions()[perm])
    assert np.array_equal(out["min"], ref["min"])
