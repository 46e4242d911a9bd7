"""The BATR topologist (Topologists/BatrTopologist.cpp): the oracle pinned by
cases the geometry decides and by invariants, and the library's host
implementation (gm_mesh_restructure on a host-only mesh) against the oracle,
bit for bit.  Both restructure with glibc's pow, exp and cos, the reference's
libm (the oracle's libm mode 0).

What pins the oracle:
* Two tets sharing a face versus the three tets around the segment joining
  their apexes (the two triangulations of five points in convex position):
  the closed-form mean ratio 12 (3|V|)^(2/3) / sum l^2 says which is better,
  and the operations must end on it from either start (a flat bipyramid
  takes the 2-3 face swap, a tall one the 3-2 edge swap); the regular
  bipyramid stays as it is.
* One tet with a single edge above max_edge is split at that edge's exact
  midpoint into two tets when both halves clear the quality floor
  (min quality / (length / max_edge)).  A single edge below min_edge merges:
  onto the fixed end when one end is fixed, at the exact midpoint when both
  are free, leaving the tets that did not hold the edge.
* The ring triangulation table: every rotation of its entries, per ring size,
  is exactly the set of triangulations of the polygon (Catalan many), and
  each entry's rotation count is the ring size over its rotational symmetry.
* On jittered cube meshes: the result is a conformal tetrahedral mesh (every
  face in one or two tets), every tet positive, Euler characteristic 1, every
  vertex used, every fixed vertex kept; when no boundary edge exceeds
  max_edge the cube's volume is kept and every boundary face lies on the
  cube's surface.
"""
import itertools

import numpy as np
import pytest

from gpumesh_amd import GpuMesh
from gpumesh_amd.meshgen import box_boundary, cube_mesh

NOSIZE = dict(min_edge=1e-9, max_edge=1e9)


def mean_ratio(p):
    """Closed-form mean ratio of tets p[..., 4, 3] (positive orientation)."""
    a, b, c, d = (p[..., k, :] for k in range(4))
    vol = np.einsum("...i,...i->...", b - a, np.cross(c - a, d - a)) / 6.0
    l2 = sum(np.sum((p[..., i, :] - p[..., j, :]) ** 2, axis=-1)
             for i, j in [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)])
    return np.sign(vol) * 12.0 * np.abs(3.0 * vol) ** (2.0 / 3.0) / l2


def orient(pos, tets):
    tets = np.array(tets, np.uint32)
    p = pos[tets.astype(np.int64)]
    vol = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    tets[vol < 0] = tets[vol < 0][:, [1, 0, 2, 3]]
    return tets


def as_sets(pos, tets):
    """Tets as sets of vertex positions (numbering-free)."""
    return sorted(tuple(sorted(tuple(pos[v]) for v in t)) for t in tets.astype(np.int64))


def both(oracle, pos, fixed, tets, K=1.0, evaluator="mean_ratio", sampler="uniform"):
    hexs = np.zeros((0, 8), np.uint32)
    om = oracle.OracleMesh(pos, fixed, tets, hexs)
    gm = GpuMesh(pos, fixed, tets, hexs, device=False)
    om.set_crew(evaluator, sampler, scaling=K)
    gm.set_crew(evaluator, sampler, scaling=K)
    return om, gm


def run_both(om, gm, **kw):
    co = om.restructure(**kw)
    cg = gm.restructure(**kw)
    assert co == cg
    po, pg = om.positions(), gm.positions()
    to, tg = om.elements()[0], gm.elements()[0]
    assert np.array_equal(po, pg) and np.array_equal(to, tg)
    assert np.array_equal(om.fixed(), _fixed_of(gm))
    return co, po, to


def _fixed_of(gm):
    # Fixed vertices come first after renumbering.
    info_fixed = gm.nverts - gm.nmovable
    out = np.zeros(gm.nverts, np.int32)
    out[:info_fixed] = 1
    return out


def bipyramid(h):
    s = np.sqrt(3.0)
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, s / 2, 0.0],
                    [0.5, s / 6, h], [0.5, s / 6, -h]])
    two = orient(pos, [[3, 0, 1, 2], [4, 0, 1, 2]])
    three = orient(pos, [[3, 0, 1, 4], [3, 1, 2, 4], [3, 2, 0, 4]])
    return pos, two, three


@pytest.mark.parametrize("h,start,best", [(0.15, "two", "three"), (0.12, "two", "three"),
                                          (1.5, "three", "two"), (1.2, "three", "two")])
def test_bipyramid_ends_on_the_better_triangulation(oracle, h, start, best):
    pos, two, three = bipyramid(h)
    q = {k: mean_ratio(pos[v.astype(np.int64)]) for k, v in (("two", two), ("three", three))}
    hm = {k: len(v) / np.sum(1.0 / v) for k, v in q.items()}
    other = "two" if best == "three" else "three"
    # The closed form decides: the better one wins on both the minimum (edge
    # swap rule) and the harmonic mean (face swap rule).
    assert q[best].min() > q[other].min() and hm[best] > hm[other]
    om, gm = both(oracle, pos, np.ones(5, np.int32), {"two": two, "three": three}[start])
    counts, p, t = run_both(om, gm, topo_passes=5, refinement_sweeps=0, **NOSIZE)
    assert as_sets(p, t) == as_sets(pos, {"two": two, "three": three}[best])
    assert counts["face_swaps" if start == "two" else "edge_swaps"] == 1
    assert np.all(mean_ratio(p[t.astype(np.int64)]) > 0)


def test_regular_bipyramid_is_kept(oracle):
    pos, two, three = bipyramid(np.sqrt(2.0 / 3.0))
    assert np.allclose(mean_ratio(pos[two.astype(np.int64)]), 1.0)
    om, gm = both(oracle, pos, np.ones(5, np.int32), two)
    counts, p, t = run_both(om, gm, topo_passes=5, refinement_sweeps=5, min_edge=0.5, max_edge=1.5)
    assert counts == {"face_swaps": 0, "edge_swaps": 0, "splits": 0, "merges": 0, "passes": 1}
    assert as_sets(p, t) == as_sets(pos, two)


@pytest.mark.parametrize("L", [2.0, 1.8, 2.4])
def test_single_long_edge_is_split_at_its_midpoint(oracle, L):
    pos = np.array([[0.0, 0.0, 0.0], [L, 0.0, 0.0], [L / 2, 0.8, 0.0], [L / 2, 0.3, 0.8]])
    tet = orient(pos, [[0, 1, 2, 3]])
    lengths = [np.linalg.norm(pos[i] - pos[j]) for i, j in itertools.combinations(range(4), 2)]
    assert sorted(lengths)[-1] == L and sorted(lengths)[-2] < 1.5
    mid = (pos[0] + pos[1]) / 2.0
    halves = orient(np.vstack([pos, mid]), [[4, 1, 2, 3], [0, 4, 2, 3]])
    allpos = np.vstack([pos, mid])
    floor = max(mean_ratio(pos[tet.astype(np.int64)]).min() / (L / 1.5), 1e-6)
    split = bool(np.all(mean_ratio(allpos[halves.astype(np.int64)]) > floor))
    assert split   # the three cases are chosen to split
    om, gm = both(oracle, pos, np.ones(4, np.int32), tet)
    counts, p, t = run_both(om, gm, topo_passes=1, refinement_sweeps=3, min_edge=0.5, max_edge=1.5)
    assert counts["splits"] == 1 and counts["merges"] == 0
    assert len(p) == 5 and len(t) == 2
    assert any(np.array_equal(x, mid) for x in p)
    assert as_sets(p, t) == as_sets(allpos, halves)
    # The new vertex is free: split(point, point) is the volume.
    assert om.fixed().sum() == 4


def _regular_tet():
    s3 = np.sqrt(3.0)
    return [np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0]), np.array([0.5, s3 / 2, 0.0]),
            np.array([0.5, s3 / 6, np.sqrt(2.0 / 3.0)])]


def test_short_edge_to_a_fixed_vertex_merges_onto_it(oracle):
    """A free vertex e 0.1 from the fixed corner d of the regular tet, joined
    to its four faces: e-d is the one edge below min_edge; merge(fixed,
    free) keeps the fixed point, so the result is the regular tet itself."""
    a, b, c, d = _regular_tet()
    cen = (a + b + c + d) / 4
    e = d + 0.1 * (cen - d) / np.linalg.norm(cen - d)
    pos = np.array([a, b, c, d, e])
    tets = orient(pos, [[4, 1, 2, 3], [0, 4, 2, 3], [0, 1, 4, 3], [0, 1, 2, 4]])
    om, gm = both(oracle, pos, np.array([1, 1, 1, 1, 0], np.int32), tets)
    counts, p, t = run_both(om, gm, topo_passes=5, refinement_sweeps=5, min_edge=0.5, max_edge=1.5)
    assert counts["merges"] == 1 and counts["splits"] == 0
    assert len(p) == 4 and om.fixed().sum() == 4
    assert as_sets(p, t) == as_sets(pos, orient(pos, [[0, 1, 2, 3]]))


@pytest.mark.parametrize("K", [1.0, 1.2])
def test_short_free_edge_merges_at_its_midpoint(oracle, K):
    """Two free vertices: e at the centroid of the regular tet (joined to its
    faces) and f 0.12 from e inside the sub-tet (e, b, c, d) (joined to that
    one's faces).  merge(free, free) is free: both end at the exact midpoint
    and the mesh is the four-tet star around it."""
    a, b, c, d = _regular_tet()
    e = (a + b + c + d) / 4
    sub = (e + b + c + d) / 4
    f = e + 0.12 * (sub - e) / np.linalg.norm(sub - e)
    pos = np.array([a, b, c, d, e, f])
    tets = orient(pos, [[0, 4, 2, 3], [0, 1, 4, 3], [0, 1, 2, 4], [5, 1, 2, 3], [4, 5, 2, 3], [4, 1, 5, 3],
                        [4, 1, 2, 5]])
    om, gm = both(oracle, pos, np.array([1, 1, 1, 1, 0, 0], np.int32), tets, K=K)
    counts, p, t = run_both(om, gm, topo_passes=5, refinement_sweeps=5, min_edge=0.5, max_edge=1.5)
    assert counts["merges"] == 1 and counts["face_swaps"] == counts["edge_swaps"] == counts["splits"] == 0
    mid = (e + f) / 2.0
    star = np.array([a, b, c, d, mid])
    assert len(p) == 5 and om.fixed().sum() == 4
    assert any(np.array_equal(x, mid) for x in p)
    assert as_sets(p, t) == as_sets(star, orient(star, [[0, 4, 2, 3], [0, 1, 4, 3], [0, 1, 2, 4], [4, 1, 2, 3]]))


def _triangulations(n):
    """All triangulations of the convex n-gon, as frozensets of triangles."""
    def rec(poly):
        if len(poly) < 3:
            return [frozenset()]
        a, b = poly[0], poly[-1]
        out = []
        for k in range(1, len(poly) - 1):
            tri = frozenset([tuple(sorted((a, poly[k], b)))])
            for left in rec(poly[: k + 1]):
                for right in rec(poly[k:]):
                    out.append(tri | left | right)
        return out
    return set(rec(list(range(n))))


def test_ring_triangulation_table_is_complete(oracle):
    from math import comb
    table = {3: [(1, [(0, 1, 2)])],
             4: [(2, [(0, 1, 2), (0, 2, 3)])],
             5: [(5, [(0, 1, 3), (0, 3, 4), (1, 2, 3)])],
             6: [(6, [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 5)]), (3, [(0, 1, 2), (0, 2, 3), (0, 3, 5), (3, 4, 5)]),
                 (3, [(0, 1, 3), (1, 2, 3), (0, 3, 4), (0, 4, 5)]), (2, [(0, 1, 2), (0, 2, 4), (2, 3, 4), (0, 4, 5)])],
             7: [(7, [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 5), (0, 5, 6)]),
                 (7, [(0, 1, 2), (0, 2, 3), (0, 3, 6), (3, 4, 6), (4, 5, 6)]),
                 (7, [(0, 1, 2), (0, 2, 3), (0, 3, 4), (0, 4, 6), (4, 5, 6)]),
                 (7, [(0, 1, 3), (1, 2, 3), (0, 3, 4), (0, 4, 5), (0, 5, 6)]),
                 (7, [(0, 1, 2), (0, 2, 4), (2, 3, 4), (0, 4, 5), (0, 5, 6)]),
                 (7, [(0, 1, 2), (0, 2, 3), (0, 3, 5), (3, 4, 5), (0, 5, 6)])]}
    for n, entries in table.items():
        want = _triangulations(n)
        assert len(want) == comb(2 * (n - 2), n - 2) // (n - 1)
        got = []
        for rot_count, tris in entries:
            rotations = [frozenset(tuple(sorted(((a + r) % n, (b + r) % n, (c + r) % n))) for a, b, c in tris)
                         for r in range(n)]
            assert len(set(rotations)) == rot_count, (n, tris)
            got += rotations[:rot_count]
        assert len(got) == len(set(got)) and set(got) == want, n


def _faces(T):
    F = np.sort(np.concatenate([T[:, [1, 2, 3]], T[:, [0, 2, 3]], T[:, [0, 1, 3]], T[:, [0, 1, 2]]]), axis=1)
    return np.unique(F, axis=0, return_counts=True)


def check_invariants(m, p, t, fixed, volume_kept):
    t = t.astype(np.int64)
    a, b, c, d = (p[t[:, k]] for k in range(4))
    vol = np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a)) / 6.0
    assert vol.min() > 0
    faces, share = _faces(t)
    assert share.max() <= 2
    edges = np.unique(np.sort(np.concatenate([t[:, [i, j]] for i, j in itertools.combinations(range(4), 2)]), axis=1),
                      axis=0)
    assert len(p) - len(edges) + len(faces) - len(t) == 1
    assert np.array_equal(np.unique(t), np.arange(len(p)))
    # Every fixed vertex is kept, at its place.
    want = {tuple(x) for x in m.pos[m.fixed.astype(bool)]}
    assert {tuple(x) for x in p[fixed.astype(bool)]} == want
    if volume_kept:
        assert abs(vol.sum() - 1.0) < 1e-12
        bnd = p[faces[share == 1]]
        on = np.zeros(len(bnd), bool)
        for ax in range(3):
            for val in (0.0, 1.0):
                on |= np.all(bnd[:, :, ax] == val, axis=1)
        assert on.all()


@pytest.mark.parametrize("n,K,jitter,seed,min_edge", [(4, 4.0, 0.3, 3, 0.5), (6, 6.0, 0.35, 3, 0.5),
                                                      (6, 4.0, 0.35, 2, 0.8), (8, 8.0, 0.35, 11, 0.5),
                                                      (6, 8.0, 0.35, 5, 0.5), (5, 9.0, 0.4, 9, 0.5)])
def test_oracle_invariants_and_library_parity_on_cubes(oracle, n, K, jitter, seed, min_edge):
    m = cube_mesh(n, "tet", jitter, seed)
    om = oracle.OracleMesh.from_cube(m)
    gm = GpuMesh.from_cube(m, device=False)
    om.set_crew("mean_ratio", "uniform", scaling=K)
    gm.set_crew("mean_ratio", "uniform", scaling=K)
    counts, p, t = run_both(om, gm, topo_passes=5, refinement_sweeps=5, min_edge=min_edge, max_edge=1.5)
    assert counts["edge_swaps"] > 0 and counts["passes"] >= 1
    # Cube boundary edges are h and h sqrt(2) long: none is split while K h sqrt(2) <= 1.5.
    check_invariants(m, p, t, om.fixed(), volume_kept=K / n * np.sqrt(2.0) <= 1.5)
    q = om.evaluate()
    assert q["min"] > 0


@pytest.mark.parametrize("evaluator,sampler,K,n", [("metric_conformity", "uniform", 6.0, 5),
                                                   ("metric_conformity", "analytic", 1.0, 5),
                                                   ("mean_ratio", "analytic", 0.6, 3),
                                                   ("mean_ratio", "analytic_reference", 5.0, 5),
                                                   ("metric_conformity", "analytic_reference", 5.0, 5)])
def test_library_parity_under_other_crews(oracle, evaluator, sampler, K, n):
    """Metric conformity and the metric-wise measurer (edge lengths by the
    adaptive walk) under the shell field and the reference's analytic field."""
    m = cube_mesh(n, "tet", 0.35, 4)
    om = oracle.OracleMesh.from_cube(m)
    gm = GpuMesh.from_cube(m, device=False)
    om.set_crew(evaluator, sampler, scaling=K)
    gm.set_crew(evaluator, sampler, scaling=K)
    counts, p, t = run_both(om, gm, topo_passes=3, refinement_sweeps=3, min_edge=0.5, max_edge=1.5)
    assert sum(counts[k] for k in ("face_swaps", "edge_swaps", "splits", "merges")) > 0
    check_invariants(m, p, t, om.fixed(), volume_kept=False)


def test_repeated_restructures_match(oracle):
    """Restructuring a restructured mesh (the second global pass starts from
    the recompiled topology and its new numbering)."""
    m = cube_mesh(6, "tet", 0.35, 7)
    om = oracle.OracleMesh.from_cube(m)
    gm = GpuMesh.from_cube(m, device=False)
    for K in (6.0, 7.0, 5.0):
        om.set_crew("mean_ratio", "uniform", scaling=K)
        gm.set_crew("mean_ratio", "uniform", scaling=K)
        run_both(om, gm, topo_passes=2, refinement_sweeps=2, min_edge=0.5, max_edge=1.5)
    assert np.array_equal(om.permutation(), gm.permutation())


def test_restructure_leaves_other_meshes_and_refuses_sliding(oracle):
    from gpumesh_amd._lib import GpuMeshError
    m = cube_mesh(3, "hex", 0.2, 1)
    gm = GpuMesh.from_cube(m, device=False)
    before = gm.positions().copy()
    assert gm.restructure() == {"face_swaps": 0, "edge_swaps": 0, "splits": 0, "merges": 0, "passes": 0}
    assert np.array_equal(gm.positions(), before)
    t = cube_mesh(3, "tet", 0.2, 1)
    bpos, bfixed, bcons, btable = box_boundary(t)
    gs = GpuMesh.constrained(bpos, bfixed, t.tets, t.hexs, bcons, btable, device=False)
    with pytest.raises(GpuMeshError, match="sliding"):
        gs.restructure()
    g = GpuMesh.from_cube(t, device=False)
    for bad in (dict(topo_passes=0), dict(min_edge=0.0), dict(min_edge=2.0, max_edge=1.0)):
        with pytest.raises(GpuMeshError):
            g.restructure(**bad)
    pos = t.pos.copy()
    pos[int(np.flatnonzero(t.fixed == 0)[0])] = np.nan
    gn = GpuMesh(pos, t.fixed, t.tets, t.hexs, device=False)
    with pytest.raises(GpuMeshError, match="not finite"):
        gn.restructure()


def test_topologist_follows_the_c_library(oracle):
    """The greedy choices hinge on last-bit differences between nearly equal
    qualities: on these cubes the oracle under the project's libm and under
    glibc (the reference's) end on different meshes, and the library's host
    topologist makes glibc's choices."""
    differ = 0
    for n, K, seed in [(8, 8.0, 11), (6, 9.0, 5), (10, 10.0, 1)]:
        m = cube_mesh(n, "tet", 0.35, seed)
        meshes = []
        for libm in (0, 1):
            om = oracle.OracleMesh.from_cube(m)
            om.set_crew("mean_ratio", "uniform", scaling=K)
            om.restructure(5, 5, 0.5, 1.5, libm=libm)
            meshes.append((om.positions(), om.elements()[0]))
        gm = GpuMesh.from_cube(m, device=False)
        gm.set_crew("mean_ratio", "uniform", scaling=K)
        gm.restructure(5, 5, 0.5, 1.5)
        assert np.array_equal(gm.positions(), meshes[0][0]) and np.array_equal(gm.elements()[0], meshes[0][1])
        differ += not (meshes[0][0].shape == meshes[1][0].shape and np.array_equal(meshes[0][1], meshes[1][1]))
    assert differ >= 1
