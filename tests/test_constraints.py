"""Sliding boundary constraints on the host: the oracle's plane and segment
snaps (PlaneConstraint::operator(), Boundaries/Constraints/FaceConstraint.cpp:180-185;
SegmentConstraint::operator(), EdgeConstraint.cpp:185-199) against what
projection geometry fixes; the product's vertex types, colour groups and
renumbering of a BoxBoundary mesh (Boundaries/BoxBoundary.cpp:7-36) against
the oracle's NodeGroups::build (DataStructures/NodeGroups.cpp:124-241); the
oracle's constrained smoothing against the box's invariants; and
gm_mesh_create_desc's input checks."""
import numpy as np
import pytest

from gpumesh_amd import box_boundary, cube_mesh
from gpumesh_amd import GpuMesh, GpuMeshError

FACE, EDGE = 1, 2


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_plane_snap_is_the_orthogonal_projection(oracle):
    rng = np.random.default_rng(1)
    for _ in range(200):
        a = rng.uniform(-3, 3, 3)
        n = rng.normal(size=3) * rng.uniform(0.1, 10)
        x = rng.uniform(-5, 5, 3)
        r = oracle.constraint_snap(FACE, a, n, x)
        u = _unit(n)
        # On the plane, reached along the normal, and a fixed point.
        assert abs(np.dot(r - a, u)) <= 1e-13
        assert np.linalg.norm(np.cross(r - x, u)) <= 1e-13
        assert np.allclose(oracle.constraint_snap(FACE, a, n, r), r, rtol=0, atol=1e-14)
        # The normal's length and sign do not matter.
        assert np.allclose(oracle.constraint_snap(FACE, a, -3.0 * n, x), r, rtol=0, atol=1e-13)


def test_plane_snap_axis_aligned_is_exact(oracle):
    """A box face leaves a snapped coordinate within half the box of the
    face exactly on it (x - s is exact there) and the other two untouched,
    bit for bit."""
    rng = np.random.default_rng(2)
    for axis in range(3):
        for s in (-1.0, 1.0):
            p = np.zeros(3)
            p[axis] = s
            for _ in range(50):
                x = rng.uniform(-1.5, 1.5, 3)
                x[axis] = s + rng.uniform(-0.5, 0.5)
                r = oracle.constraint_snap(FACE, p, p, x)
                assert r[axis] == s
                assert all(r[i] == x[i] for i in range(3) if i != axis)


def test_segment_snap_projects_and_clamps(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.uniform(-3, 3, 3)
        b = a + rng.normal(size=3) * rng.uniform(0.1, 5)
        u = _unit(b - a)
        L = np.linalg.norm(b - a)
        x = rng.uniform(-6, 6, 3)
        r = oracle.constraint_snap(EDGE, a, b, x)
        t = np.dot(x - a, u)
        if t < 0:
            assert np.array_equal(r, a)           # clamped to the first end, exactly
        elif t > L:
            assert np.array_equal(r, b)           # and to the second
        else:
            assert np.linalg.norm(np.cross(r - a, u)) <= 1e-12
            assert abs(np.dot(x - r, u)) <= 1e-12  # the move is orthogonal to the segment
            assert np.allclose(r, a + t * u, rtol=0, atol=1e-12)


def test_segment_snap_box_edge_is_exact(oracle):
    rng = np.random.default_rng(4)
    a, b = np.array([-1.0, 1.0, -1.0]), np.array([1.0, 1.0, -1.0])
    for _ in range(100):
        x = np.array([rng.uniform(-0.99, 0.99), rng.uniform(0.9, 1.1), rng.uniform(-1.1, -0.9)])
        r = oracle.constraint_snap(EDGE, a, b, x)
        assert r[0] == x[0] and r[1] == 1.0 and r[2] == -1.0


def _box(kind, n, jitter=0.25, seed=5):
    m = cube_mesh(n, kind, jitter, seed=seed)
    return (m,) + box_boundary(m, 0.2, seed)


def _classes(pos, fixed, cons, constraints):
    kinds = np.array([c[0] for c in constraints])
    face = (cons >= 0) & (kinds[np.maximum(cons, 0)] == FACE)
    edge = (cons >= 0) & (kinds[np.maximum(cons, 0)] == EDGE)
    return fixed.astype(bool), face, edge


@pytest.mark.parametrize("kind,n", [("tet", 3), ("tet", 6), ("hex", 4), ("hex", 7)])
def test_box_topology_matches_oracle(oracle, kind, n):
    m, pos, fixed, cons, constraints = _box(kind, n)
    om = oracle.OracleMesh(pos, fixed, m.tets, m.hexs, constraints=constraints, constraint=cons)
    gm = GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, constraints, device=False)
    assert np.array_equal(om.permutation(), gm.permutation())
    eo1, er1, no1, nb1 = om.topology()
    eo2, er2, no2, nb2 = gm.topology()
    assert np.array_equal(eo1, eo2.astype(np.int64)) and np.array_equal(no1, no2.astype(np.int64))
    assert np.array_equal(nb1, nb2)
    t1, h1 = om.elements()
    t2, h2 = gm.elements()
    assert np.array_equal(t1, t2) and np.array_equal(h1, h2)
    groups, tr = om.groups()
    off, nodes = gm.groups()
    assert gm.ngroups == len(groups)
    # Type blocks: fixed = the corners, boundary = face and edge vertices.
    isfixed, face, edge = _classes(pos, fixed, cons, constraints)
    assert tr[0, 1] - tr[0, 0] == isfixed.sum() == 8
    assert tr[1, 1] - tr[1, 0] == face.sum() + edge.sum()
    assert tr[2, 1] > tr[2, 0] and (tr[3, 1] > tr[3, 0]) == (n >= 4)
    # Each colour group: its boundary, subsurface and interior runs, in that order.
    for g in range(gm.ngroups):
        want = np.concatenate([np.arange(groups[g, 0], groups[g, 1]), np.arange(groups[g, 2], groups[g, 3]),
                               np.arange(groups[g, 4], groups[g, 5])])
        assert np.array_equal(nodes[off[g]:off[g + 1]], want)


def test_shuffled_box_topology_matches_oracle(oracle):
    m, pos, fixed, cons, constraints = _box("tet", 5)
    rng = np.random.default_rng(9)
    perm = rng.permutation(len(pos))
    inv = np.argsort(perm)
    pos, fixed, cons = pos[perm], fixed[perm], cons[perm]
    tets = inv[m.tets][rng.permutation(len(m.tets))].astype(np.uint32)
    om = oracle.OracleMesh(pos, fixed, tets, m.hexs, constraints=constraints, constraint=cons)
    gm = GpuMesh.constrained(pos, fixed, tets, m.hexs, cons, constraints, device=False)
    assert np.array_equal(om.permutation(), gm.permutation())
    groups, _ = om.groups()
    off, nodes = gm.groups()
    for g in range(gm.ngroups):
        want = np.concatenate([np.arange(groups[g, 2 * t], groups[g, 2 * t + 1]) for t in range(3)])
        assert np.array_equal(nodes[off[g]:off[g + 1]], want)


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_oracle_box_smoothing_slides_on_the_box(oracle, smoother, kind):
    """After constrained sweeps every corner is where it was, every face
    vertex lies exactly on its face and every edge vertex exactly on its
    edge, inside the box; face and edge vertices did move; the mesh
    improved."""
    m, pos, fixed, cons, constraints = _box(kind, 5)
    om = oracle.OracleMesh(pos, fixed, m.tets, m.hexs, constraints=constraints, constraint=cons)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    st = om.smooth(smoother, 2, threads=2, spawn_coeff=coeff)
    out = om.positions()[om.permutation()]          # back to the input numbering
    isfixed, face, edge = _classes(pos, fixed, cons, constraints)
    assert np.array_equal(out[isfixed], pos[isfixed])
    on = np.abs(out) == 1.0
    assert np.all(on[face].sum(axis=1) >= 1) and np.all(on[edge].sum(axis=1) >= 2)
    for v in np.where(face | edge)[0]:
        kind_, a, b = constraints[cons[v]]
        if kind_ == FACE:
            axis = int(np.argmax(np.abs(b)))
            assert out[v, axis] == a[axis]
        else:
            for i in range(3):
                if a[i] == b[i]:
                    assert out[v, i] == a[i]
    assert np.all(np.abs(out) <= 1.0)
    moved = np.any(out != pos, axis=1)
    assert moved[face].sum() > face.sum() // 2 and moved[edge].sum() > 0
    assert st["mean"][-1] > st["mean"][0]


def test_snapping_changes_the_result(oracle):
    """The same mesh with every boundary vertex fixed ends elsewhere: the
    constrained sweep is not the fixed-boundary one."""
    m, pos, fixed, cons, constraints = _box("tet", 4)
    a = oracle.OracleMesh(pos, fixed, m.tets, m.hexs, constraints=constraints, constraint=cons)
    b = oracle.OracleMesh(pos, (cons >= 0) | (fixed > 0), m.tets, m.hexs)
    a.smooth("nm", 1)
    b.smooth("nm", 1)
    pa = a.positions()[a.permutation()]
    pb = b.positions()[b.permutation()]
    assert not np.array_equal(pa, pb)


def test_create_desc_rejects_bad_input():
    m, pos, fixed, cons, constraints = _box("tet", 2)
    bad = list(constraints)
    bad[0] = (7, bad[0][1], bad[0][2])
    with pytest.raises(GpuMeshError, match="unknown kind"):
        GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, bad, device=False)
    bad = list(constraints)
    bad[0] = (FACE, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0))
    with pytest.raises(GpuMeshError, match="zero or non-finite"):
        GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, bad, device=False)
    bad = list(constraints)
    bad[13] = (EDGE, (1.0, 1.0, 1.0), (1.0, 1.0, 1.0))
    with pytest.raises(GpuMeshError, match="zero or non-finite"):
        GpuMesh.constrained(pos, fixed, m.tets, m.hexs, cons, bad, device=False)
    c2 = cons.copy()
    c2[np.where(cons >= 0)[0][0]] = len(constraints)
    with pytest.raises(GpuMeshError, match="names constraint"):
        GpuMesh.constrained(pos, fixed, m.tets, m.hexs, c2, constraints, device=False)
    c2[np.where(cons >= 0)[0][0]] = -2
    with pytest.raises(GpuMeshError, match="names constraint"):
        GpuMesh.constrained(pos, fixed, m.tets, m.hexs, c2, constraints, device=False)


def test_fixed_and_sliding_is_refused():
    """A vertex is either fixed (the reference's VertexConstraint) or slides
    on one face or edge constraint (AbstractConstraint.h:65-73): naming a
    constraint for a fixed vertex is refused."""
    m, pos, fixed, cons, constraints = _box("tet", 3)
    c2 = cons.copy()
    c2[fixed > 0] = 0
    with pytest.raises(GpuMeshError, match="both fixed and sliding"):
        GpuMesh.constrained(pos, fixed, m.tets, m.hexs, c2, constraints, device=False)


def test_no_constraints_equals_plain_create(oracle):
    """A descriptor without constraints builds what gm_mesh_create builds."""
    m = cube_mesh(5, "hex", 0.3, seed=2)
    a = GpuMesh.from_cube(m, device=False)
    b = GpuMesh.constrained(m.pos, m.fixed, m.tets, m.hexs, -np.ones(len(m.pos), np.int32), [], device=False)
    assert np.array_equal(a.permutation(), b.permutation())
    assert all(np.array_equal(x, y) for x, y in zip(a.topology(), b.topology()))
    assert all(np.array_equal(x, y) for x, y in zip(a.groups(), b.groups()))
