"""Product topology and colouring (gm_topology.cpp, host only) against the
oracle's literal restatement of Mesh::compileNeighborhoods and
NodeGroups::build: bit-exact permutation, patch lists, neighbour lists,
renumbered elements and colour groups."""
import numpy as np
import pytest

from gpumesh_amd import cube_mesh
from gpumesh_amd.mesh import GpuMesh


def _compare(oracle, pos, fixed, tets, hexs):
    om = oracle.OracleMesh(pos, fixed, tets, hexs)
    gm = GpuMesh(pos, fixed, tets, hexs, device=False)
    assert np.array_equal(om.permutation(), gm.permutation())
    eo1, er1, no1, nb1 = om.topology()
    eo2, er2, no2, nb2 = gm.topology()
    # oracle packs (id << 5) | (type << 3) | local; product (id << 4) | (is_hex << 3) | local
    er1c = ((er1 >> 5) << 4) | ((((er1 >> 3) & 3) == 3).astype(np.uint64) << 3) | (er1 & 7)
    assert np.array_equal(eo1, eo2.astype(np.int64))
    assert np.array_equal(er1c.astype(np.uint32), er2)
    assert np.array_equal(no1, no2.astype(np.int64))
    assert np.array_equal(nb1, nb2)
    t1, h1 = om.elements()
    t2, h2 = gm.elements()
    assert np.array_equal(t1, t2) and np.array_equal(h1, h2)
    groups, tr = om.groups()
    off, nodes = gm.groups()
    assert gm.ngroups == len(groups)
    nfixed = tr[0, 1]
    for g in range(gm.ngroups):
        if groups[g, 5] > groups[g, 4]:
            assert off[g] + nfixed == groups[g, 4] and off[g + 1] + nfixed == groups[g, 5]
        else:
            assert off[g] == off[g + 1]
    return om, gm


@pytest.mark.parametrize("kind,n", [("tet", 1), ("tet", 5), ("tet", 12), ("hex", 1), ("hex", 7), ("hex", 15)])
def test_cube_meshes(oracle, kind, n):
    m = cube_mesh(n, kind, 0.3, seed=n)
    _compare(oracle, m.pos, m.fixed, m.tets, m.hexs)


def test_shuffled_numbering_and_mixed_elements(oracle):
    """Random vertex ids, random element order, tets and hexes together,
    free boundary vertices: exercises BFS restarts, ties in the sort and
    colours used only by fixed vertices."""
    rng = np.random.default_rng(11)
    t = cube_mesh(6, "tet", 0.0, seed=1)
    h = cube_mesh(6, "hex", 0.0, seed=1)
    cx_t = t.pos[t.tets].mean(axis=1)[:, 0]
    cx_h = h.pos[h.hexs].mean(axis=1)[:, 0]
    tets = t.tets[cx_t > 0.5]
    hexs = h.hexs[cx_h < 0.5]
    perm = rng.permutation(t.nverts)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(t.nverts)
    pos = t.pos[perm]
    fixed = (rng.uniform(size=t.nverts) < 0.2).astype(np.int32)
    tets = inv[tets][rng.permutation(len(tets))].astype(np.uint32)
    hexs = inv[hexs][rng.permutation(len(hexs))].astype(np.uint32)
    _compare(oracle, pos, fixed, tets, hexs)


def test_shuffled_tet_only(oracle):
    """Tetrahedra only (the colouring walks neighbour lists): random vertex
    ids, random element order, random fixed vertices."""
    rng = np.random.default_rng(5)
    t = cube_mesh(7, "tet", 0.2, seed=4)
    perm = rng.permutation(t.nverts)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(t.nverts)
    fixed = (rng.uniform(size=t.nverts) < 0.15).astype(np.int32)
    tets = inv[t.tets][rng.permutation(len(t.tets))].astype(np.uint32)
    _compare(oracle, t.pos[perm], fixed, tets, np.zeros((0, 8), np.uint32))


def test_disconnected_components(oracle):
    a = cube_mesh(3, "tet", 0.1, seed=2)
    b = cube_mesh(2, "hex", 0.1, seed=3)
    pos = np.concatenate([a.pos, b.pos + 5.0])
    fixed = np.concatenate([a.fixed, b.fixed])
    hexs = b.hexs + a.nverts
    _compare(oracle, pos, fixed, a.tets, hexs.astype(np.uint32))


def test_large_kuhn_cube_groups(oracle):
    m = cube_mesh(40, "tet", 0.35, seed=2)
    om, gm = _compare(oracle, m.pos, m.fixed, m.tets, m.hexs)
    assert gm.max_patch == 24 and gm.nmovable == 39 ** 3


def test_topology_cache_hits_are_exact(oracle):
    """With the topology cache on, meshes that share connectivity reuse one
    build; a hit must give exactly the topology of a fresh build, and a mesh
    whose connectivity or fixed flags differ must miss."""
    from gpumesh_amd._lib import check, lib
    import ctypes

    def stats():
        out = (ctypes.c_int64 * 4)()
        check(lib().gm_topology_cache_stats(out), "stats")
        return list(out)

    check(lib().gm_set_topology_cache(1 << 30), "cache on")
    try:
        a = cube_mesh(9, "tet", 0.3, seed=1)
        b = cube_mesh(9, "tet", 0.3, seed=2)        # same connectivity, other positions
        s0 = stats()
        ga = GpuMesh.from_cube(a, device=False)
        s1 = stats()
        _, gb = _compare(oracle, b.pos, b.fixed, b.tets, b.hexs)
        s2 = stats()
        assert s1[1] == s0[1] + 1 and s2[0] == s1[0] + 1       # a misses, b hits
        assert np.array_equal(ga.permutation(), gb.permutation())
        fixed2 = b.fixed.copy()
        fixed2[np.nonzero(fixed2 == 0)[0][0]] = 1               # one more fixed vertex: must miss
        _compare(oracle, b.pos, fixed2, b.tets, b.hexs)
        s3 = stats()
        assert s3[1] == s2[1] + 1
        tets2 = b.tets.copy()
        tets2[[0, 1]] = tets2[[1, 0]]                           # element order matters: must miss
        _compare(oracle, b.pos, b.fixed, tets2, b.hexs)
        assert stats()[1] == s3[1] + 1
    finally:
        check(lib().gm_set_topology_cache(0), "cache off")
    assert stats()[2] == 0


def test_topology_prepare_host_only():
    """gm_topology_prepare: refused with the cache off, refused for a
    topology larger than the capacity, and a cache hit for the next mesh of
    that shape otherwise."""
    from gpumesh_amd import GpuMeshError, prepare_topology, set_topology_cache, topology_cache_stats
    m = cube_mesh(6, "hex", 0.2, seed=4)
    with pytest.raises(GpuMeshError, match="cache is off"):
        prepare_topology(m.pos, m.fixed, m.tets, m.hexs, device=False)
    set_topology_cache(4096)
    try:
        with pytest.raises(GpuMeshError, match="capacity"):
            prepare_topology(m.pos, m.fixed, m.tets, m.hexs, device=False)
        set_topology_cache(1 << 28)
        prepare_topology(m.pos, m.fixed, m.tets, m.hexs, device=False)
        before = topology_cache_stats()
        assert before["entries"] == 1
        other = cube_mesh(6, "hex", 0.3, seed=9)
        g = GpuMesh.from_cube(other, device=False)
        assert topology_cache_stats()["hits"] == before["hits"] + 1
        ref = GpuMesh(other.pos, other.fixed, other.tets, other.hexs, device=False)
        assert np.array_equal(g.permutation(), ref.permutation())
    finally:
        set_topology_cache(0)


def _compare_with_prisms(oracle, pos, fixed, tets, hexs, pris):
    om = oracle.OracleMesh(pos, fixed, tets, hexs, pris=pris)
    gm = GpuMesh(pos, fixed, tets, hexs, device=False, pris=pris)
    assert gm.npris == len(pris)
    assert np.array_equal(om.permutation(), gm.permutation())
    eo1, er1, no1, nb1 = om.topology()
    eo2, er2, no2, nb2 = gm.topology()
    # oracle: (id << 5) | (type << 3) | local with type 0 tet, 2 prism, 3 hex;
    # product: (record id << 4) | (is_record << 3) | local, prism p = record H + p.
    typ = (er1 >> np.uint64(3)) & np.uint64(3)
    oid = er1 >> np.uint64(5)
    rid = np.where(typ == 2, oid + np.uint64(len(hexs)), oid)
    er1c = (rid << np.uint64(4)) | ((typ >= 2).astype(np.uint64) << np.uint64(3)) | (er1 & np.uint64(7))
    assert np.array_equal(eo1, eo2.astype(np.int64))
    assert np.array_equal(er1c.astype(np.uint32), er2)
    assert np.array_equal(no1, no2.astype(np.int64))
    assert np.array_equal(nb1, nb2)
    t1, h1 = om.elements()
    t2, h2 = gm.elements()
    assert np.array_equal(t1, t2) and np.array_equal(h1, h2)
    assert np.array_equal(om.prisms(), gm.prisms())
    groups, tr = om.groups()
    off, nodes = gm.groups()
    assert gm.ngroups == len(groups)
    nfixed = tr[0, 1]
    for g in range(gm.ngroups):
        if groups[g, 5] > groups[g, 4]:
            assert off[g] + nfixed == groups[g, 4] and off[g + 1] + nfixed == groups[g, 5]
        else:
            assert off[g] == off[g + 1]


@pytest.mark.parametrize("layout,n,seed", [("prism", 1, 1), ("prism", 5, 2), ("mixed", 6, 3), ("mixed", 9, 4)])
def test_prism_meshes(oracle, layout, n, seed):
    from gpumesh_amd.meshgen import element_cube_mesh
    m = element_cube_mesh(n, layout, 0.3, seed)
    _compare_with_prisms(oracle, m.pos, m.fixed, m.tets, m.hexs, m.pris)


def test_shuffled_prism_mesh(oracle):
    """Random vertex ids, random element order and free boundary vertices
    with prisms between the tetrahedra and hexahedra."""
    from gpumesh_amd.meshgen import element_cube_mesh
    rng = np.random.default_rng(5)
    m = element_cube_mesh(6, "mixed", 0.2, 8)
    perm = rng.permutation(m.nverts)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.nverts)
    pos = m.pos[perm]
    fixed = (rng.uniform(size=m.nverts) < 0.2).astype(np.int32)
    tets = inv[m.tets][rng.permutation(len(m.tets))].astype(np.uint32)
    hexs = inv[m.hexs][rng.permutation(len(m.hexs))].astype(np.uint32)
    pris = inv[m.pris][rng.permutation(len(m.pris))].astype(np.uint32)
    _compare_with_prisms(oracle, pos, fixed, tets, hexs, pris)
