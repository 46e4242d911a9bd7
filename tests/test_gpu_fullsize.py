"""Full-size checks on the 1M-vertex Kuhn cube (100^3 cells, jitter 0.35 h,
seed 2), where a whole-mesh CPU oracle run would take minutes.

Vertices of one colour group share no element, so a vertex's move depends
only on the positions at the start of its group's launch.  The GPU runs the
groups before g, the oracle is given that state and re-smooths a random
sample of group g's vertices one by one, and the GPU's group-g launch must
produce the same bits for every sampled vertex.  The rest of the file checks
size-independent properties: kernel shapes agree bit for bit, the
evaluation counts match, no element inverts, quality improves.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh

pytestmark = pytest.mark.gpu

N_SAMPLE = 1500


@pytest.fixture(scope="module")
def c2():
    m = cube_mesh(100, "tet", 0.35, seed=2)
    gm = GpuMesh.from_cube(m)
    return m, gm, gm.positions()


@pytest.fixture(scope="module")
def c2_oracle(oracle, c2):
    m, gm, start = c2
    om = oracle.OracleMesh.from_cube(m)
    assert np.array_equal(om.positions(), start)
    assert np.array_equal(om.permutation(), gm.permutation())
    return om


@pytest.mark.parametrize("smoother,coeff", [("gd", 0.05), ("spawn", 0.25), ("nm", 0.05)])
@pytest.mark.parametrize("group", [0, 5])
def test_group_replay_bitexact(oracle, c2, c2_oracle, smoother, coeff, group):
    m, gm, start = c2
    om = c2_oracle
    gm.set_positions(start)
    if group > 0:
        gm.smooth_groups(smoother, 0, group, "fp64", "wave", coeff)
    before = gm.positions()
    counts = gm.smooth_groups(smoother, group, group + 1, "fp64", "wave", coeff)
    after = gm.positions()
    off, nodes = gm.groups()
    members = nodes[off[group]:off[group + 1]]
    rng = np.random.default_rng(100 + group)
    sample = np.sort(rng.choice(members, size=min(N_SAMPLE, members.size), replace=False)).astype(np.uint32)
    om.set_positions(before)
    ref_counts = om.smooth_vertices(smoother, sample, spawn_coeff=coeff)
    ref = om.positions()
    assert np.array_equal(after[sample], ref[sample])
    assert np.array_equal(counts[sample], ref_counts.astype(np.int32))
    # Vertices outside group g did not move.
    outside = np.ones(m.nverts, bool)
    outside[members] = False
    assert np.array_equal(after[outside], before[outside])


@pytest.mark.parametrize("smoother,coeff", [("gd", 0.05), ("spawn", 0.25), ("nm", 0.05)])
def test_group_replay_against_glibc_oracle(oracle, c2, c2_oracle, smoother, coeff):
    """The exact FP64 kernels (the project's correctly rounded pow/cbrt/exp)
    against the oracle calling glibc, the reference's libm: group 3's sampled
    vertices within 1e-9 h, their patch qualities within 1e-12."""
    m, gm, start = c2
    om = c2_oracle
    group = 3
    gm.set_positions(start)
    gm.smooth_groups(smoother, 0, group, "fp64", "wave", coeff)
    before = gm.positions()
    gm.smooth_groups(smoother, group, group + 1, "fp64", "wave", coeff)
    after = gm.positions()
    off, nodes = gm.groups()
    members = nodes[off[group]:off[group + 1]]
    sample = np.sort(np.random.default_rng(7).choice(members, size=N_SAMPLE, replace=False)).astype(np.uint32)
    om.set_positions(before)
    oracle.set_libm(0)
    try:
        om.smooth_vertices(smoother, sample, spawn_coeff=coeff)
        ref = om.positions()
        q_ref = np.array([om.patch_quality(int(v)) for v in sample])
        om.set_positions(after)
        q_gpu = np.array([om.patch_quality(int(v)) for v in sample])
    finally:
        oracle.set_libm(1)
    assert np.max(np.abs(after[sample] - ref[sample])) <= 1e-9 * 0.01
    assert np.max(np.abs(q_gpu - q_ref)) <= 1e-12


def test_kernel_shapes_agree_fullsize(c2):
    m, gm, start = c2
    res = {}
    for kernel in ("plain", "wave"):
        gm.set_positions(start)
        r = gm.smooth("gd", 2, precision="fp64", kernel=kernel)
        res[kernel] = (gm.positions(), r)
    assert np.array_equal(res["plain"][0], res["wave"][0])
    assert np.array_equal(res["plain"][1]["eval_counts"], res["wave"][1]["eval_counts"])
    assert np.array_equal(res["plain"][1]["min"], res["wave"][1]["min"])


@pytest.mark.parametrize("smoother", ["gd", "getme", "spawn"])
def test_mixed_precision_fullsize_quality(c2, smoother):
    """FP32 bulk + FP64 polish vs FP64 on the full mesh: final min/mean
    within the stated tolerances, no inverted element, mean improves."""
    m, gm, start = c2
    coeff = 0.25 if smoother == "spawn" else 0.05
    gm.set_positions(start)
    r64 = gm.smooth(smoother, 5, precision="fp64", kernel="wave", spawn_coeff=coeff)
    p64 = gm.positions()
    gm.set_positions(start)
    r32 = gm.smooth(smoother, 5, precision="fp32_polish", kernel="wave", spawn_coeff=coeff)
    p32 = gm.positions()
    h = 1.0 / 100
    assert r32["min"][-1] > 0 and r64["min"][-1] > 0
    assert r32["mean"][-1] > r32["mean"][0]
    assert abs(r32["mean"][-1] - r64["mean"][-1]) < 2e-4
    assert np.max(np.abs(p32 - p64)) < 0.5 * h
