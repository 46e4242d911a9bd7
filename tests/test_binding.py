"""The drop-in boundary in the reference's own numbering: gm_mesh_create_grouped
takes the mesh as NodeGroups has clustered it (positions, elements and the
parallel groups' undispatchedNodes), keeps that numbering and checks the
groups; include/gpumesh_binding.hpp, compiled with g++ against stand-ins of
the reference's types, maps a Mesh onto it (tests/host/binding_test.cpp)."""
import os
import subprocess

import numpy as np
import pytest

from gpumesh_amd import GpuMeshError, box_boundary, cube_mesh
from gpumesh_amd.mesh import GpuMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "host")
EXE = os.path.join(HOST, "_build", "binding_test")


def clustered(oracle, kind, n, jitter=0.3, seed=3):
    """The mesh as the reference holds it after NodeGroups::build: the
    oracle's renumbered positions, elements and colour groups."""
    m = cube_mesh(n, kind, jitter, seed=seed)
    om = oracle.OracleMesh.from_cube(m)
    groups, tr = om.groups()
    pos = om.positions()
    perm = om.permutation()
    fixed = np.empty(m.nverts, np.int32)
    fixed[perm] = m.fixed
    tets, hexs = om.elements()
    gnodes = np.concatenate([np.arange(g[4], g[5], dtype=np.uint32) for g in groups]) if len(groups) else \
        np.zeros(0, np.uint32)
    goff = np.concatenate([[0], np.cumsum([g[5] - g[4] for g in groups])]).astype(np.int64)
    return m, om, pos, fixed, tets, hexs, goff, gnodes


def clustered_box(oracle, kind, n, seed=5):
    """A BoxBoundary mesh (faces and edges sliding) as the reference holds it
    after NodeGroups::build: renumbered type-major, each parallel group's
    undispatchedNodes its boundary, subsurface and interior ranges."""
    m = cube_mesh(n, kind, 0.25, seed=seed)
    pos0, fixed0, cons0, constraints = box_boundary(m, 0.2, seed)
    om = oracle.OracleMesh(pos0, fixed0, m.tets, m.hexs, constraints=constraints, constraint=cons0)
    groups, _ = om.groups()
    perm = om.permutation()
    fixed = np.empty(m.nverts, np.int32)
    fixed[perm] = fixed0
    cons = np.empty(m.nverts, np.int32)
    cons[perm] = cons0
    tets, hexs = om.elements()
    runs = [np.concatenate([np.arange(g[2 * t], g[2 * t + 1], dtype=np.uint32) for t in range(3)]) for g in groups]
    gnodes = np.concatenate(runs).astype(np.uint32)
    goff = np.concatenate([[0], np.cumsum([len(r) for r in runs])]).astype(np.int64)
    ctab = np.array([[k] + list(a) + list(b) for k, a, b in constraints], np.float64)
    return m, om, om.positions(), fixed, tets, hexs, goff, gnodes, cons, ctab


def write_inputs(d, pos, fixed, tets, hexs, goff, gnodes, cons=None, ctab=None):
    arrays = [("pos", pos, np.float64), ("fixed", fixed, np.int32), ("tets", tets, np.uint32),
              ("hexs", hexs, np.uint32), ("goff", goff, np.int64), ("gnodes", gnodes, np.uint32)]
    if cons is not None:
        arrays += [("cons", cons, np.int32), ("ctab", ctab, np.float64)]
    for name, a, dt in arrays:
        np.ascontiguousarray(a, dt).tofile(os.path.join(d, name + ".bin"))


@pytest.fixture(scope="module")
def exe():
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    return EXE


@pytest.mark.parametrize("kind,n", [("tet", 6), ("hex", 5)])
def test_binding_keeps_the_reference_numbering(oracle, exe, tmp_path, kind, n):
    m, om, pos, fixed, tets, hexs, goff, gnodes = clustered(oracle, kind, n)
    d = str(tmp_path)
    write_inputs(d, pos, fixed, tets, hexs, goff, gnodes)
    r = subprocess.run([exe, "host", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "host ok" in r.stdout
    if kind == "tet":
        assert "refused" in r.stdout and "share" in r.stdout
    perm = np.fromfile(os.path.join(d, "out_perm.bin"), np.int64)
    assert np.array_equal(perm, np.arange(m.nverts))
    assert np.array_equal(np.fromfile(os.path.join(d, "out_tets.bin"), np.uint32), tets.reshape(-1))
    assert np.array_equal(np.fromfile(os.path.join(d, "out_hexs.bin"), np.uint32), hexs.reshape(-1))
    assert np.array_equal(np.fromfile(os.path.join(d, "out_goff.bin"), np.int64), goff)
    assert np.array_equal(np.fromfile(os.path.join(d, "out_gnodes.bin"), np.uint32), gnodes)


@pytest.mark.parametrize("kind,n", [("tet", 5), ("hex", 5)])
def test_binding_describes_sliding_vertices(oracle, exe, tmp_path, kind, n):
    """Stand-in PlaneConstraint and SegmentConstraint objects described to
    the binding by their geometry: the handle keeps the numbering and the
    groups, and the same mesh without a description is refused."""
    m, om, pos, fixed, tets, hexs, goff, gnodes, cons, ctab = clustered_box(oracle, kind, n)
    d = str(tmp_path)
    write_inputs(d, pos, fixed, tets, hexs, goff, gnodes, cons, ctab)
    r = subprocess.run([exe, "host", d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "host ok" in r.stdout and "undescribed refused" in r.stdout and "no plane or segment" in r.stdout
    assert np.array_equal(np.fromfile(os.path.join(d, "out_perm.bin"), np.int64), np.arange(m.nverts))
    assert np.array_equal(np.fromfile(os.path.join(d, "out_goff.bin"), np.int64), goff)
    assert np.array_equal(np.fromfile(os.path.join(d, "out_gnodes.bin"), np.uint32), gnodes)


@pytest.mark.parametrize("n,K", [(6, 6.0), (7, 8.0)])
def test_binding_restructure_hands_back_the_mesh_before_compile(oracle, exe, tmp_path, n, K):
    """Session::restructure on a mesh in the reference's numbering: the mesh it
    writes back, compiled as the reference compiles it (the oracle's
    compileNeighborhoods and NodeGroups build), is the mesh the oracle's
    BATR restructure ends on, and the handle's numbering."""
    m, om, pos, fixed, tets, hexs, goff, gnodes = clustered(oracle, "tet", n)
    d = str(tmp_path)
    write_inputs(d, pos, fixed, tets, hexs, goff, gnodes)
    r = subprocess.run([exe, "restructure", d, str(K)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    counts = np.fromfile(os.path.join(d, "out_counts.bin"), np.int64)
    back_pos = np.fromfile(os.path.join(d, "out_pos.bin"), np.float64).reshape(-1, 3)
    back_fixed = np.fromfile(os.path.join(d, "out_fixed.bin"), np.int32)
    back_tets = np.fromfile(os.path.join(d, "out_tets.bin"), np.uint32).reshape(-1, 4)
    ref = om   # the mesh the reference holds (creating one from the arrays would cluster it again)
    ref.set_crew("mean_ratio", "uniform", scaling=K)
    want = ref.restructure(5, 5, 0.5, 1.5)
    assert list(counts) == [want[k] for k in ("face_swaps", "edge_swaps", "splits", "merges", "passes")]
    compiled = oracle.OracleMesh(back_pos, back_fixed, back_tets, np.zeros((0, 8), np.uint32))
    assert np.array_equal(compiled.positions(), ref.positions())
    assert np.array_equal(compiled.elements()[0], ref.elements()[0])
    assert np.array_equal(compiled.fixed(), ref.fixed())


def test_grouped_topology_equals_the_clustered_build(oracle):
    """A mesh in its original numbering with the reference's colouring mapped
    back to it: patch and neighbour lists are the renumbering build's, mapped."""
    m = cube_mesh(7, "tet", 0.3, seed=4)
    ref = GpuMesh.from_cube(m, device=False)
    perm = ref.permutation()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.nverts)
    off, nodes = ref.groups()
    g = GpuMesh.grouped(m.pos, m.fixed, m.tets, m.hexs, off, inv[nodes], device=False)
    assert np.array_equal(g.permutation(), np.arange(m.nverts))
    eo1, er1, no1, nb1 = ref.topology()
    eo2, er2, no2, nb2 = g.topology()
    for v in range(m.nverts):
        p = perm[v]
        assert np.array_equal(er2[eo2[v]:eo2[v + 1]], er1[eo1[p]:eo1[p + 1]])
        assert np.array_equal(nb2[no2[v]:no2[v + 1]], inv[nb1[no1[p]:no1[p + 1]]])
    assert g.ngroups == ref.ngroups and g.max_patch == ref.max_patch


def test_grouped_rejects_bad_groups():
    m = cube_mesh(4, "tet", 0.2, seed=1)
    ref = GpuMesh.from_cube(m, device=False)
    perm = ref.permutation()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.nverts)
    off, nodes = ref.groups()
    nodes = inv[nodes].astype(np.uint32)
    fixed_v = int(np.where(m.fixed)[0][0])
    cases = {}
    # Two element-mates in one group.
    t0 = m.tets[np.all(m.fixed[m.tets] == 0, axis=1)][0] if np.any(np.all(m.fixed[m.tets] == 0, axis=1)) else None
    a = int(nodes[0])
    mate = next(int(x) for t in m.tets if a in t for x in t if x != a and not m.fixed[x])
    bad = nodes.copy()
    j = int(np.where(bad == mate)[0][0])
    i = int(np.where(bad == a)[0][0])
    # Move the mate into a's group by swapping it with another member of that group.
    gi = int(np.searchsorted(off, i, side="right") - 1)
    k = next(x for x in range(off[gi], off[gi + 1]) if x != i)
    bad[k], bad[j] = bad[j], bad[k]
    cases["shared element"] = (off, bad)
    dup = nodes.copy()
    dup[1] = dup[0]
    cases["vertex twice"] = (off, dup)
    fx = nodes.copy()
    fx[0] = fixed_v
    cases["fixed vertex"] = (off, fx)
    cases["missing vertex"] = (np.r_[off[:-1], off[-1] - 1], nodes[:-1])
    for name, (o, nd) in cases.items():
        with pytest.raises(GpuMeshError) as e:
            GpuMesh.grouped(m.pos, m.fixed, m.tets, m.hexs, o, nd, device=False)
        assert e.value.status == 1, name
