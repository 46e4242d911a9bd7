"""Full-size checks on the anisotropic hex workload: 64^3 hexahedra (274,625
vertices, jitter 0.3 h, seed 3), metric-conformity evaluator with the
metric-wise measurer under the spherical-shell metric (K = 64), evaluated
analytically and through a 64^3 FP32 texture.

As for the tet cube, one colour group's vertices are independent, so the GPU
runs the groups before g in FP64, the oracle re-smooths a random sample of
group g's vertices from that state, and the GPU's group-g launch must give the
same bits and evaluation counts for every sampled vertex.  The 15-pass
Nelder-Mead run of the workload is checked for size-independent properties:
no inversion, quality improves, and the FP32 + FP64-polish result agrees
with the FP64 one.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh
from gpumesh_amd.samplers import shell_texture

pytestmark = pytest.mark.gpu

N = 64
K = 64.0
N_SAMPLE = 1000


@pytest.fixture(scope="module")
def c3():
    m = cube_mesh(N, "hex", 0.3, seed=3)
    return m, shell_texture(m.pos, N, K)


def _crew(obj, sampler, tex):
    obj.set_crew("metric_conformity", sampler, scaling=K, texture=tex if sampler == "texture" else None)


@pytest.mark.parametrize("sampler", ["analytic", "texture"])
@pytest.mark.parametrize("smoother", ["nm", "gd"])
@pytest.mark.parametrize("group", [0, 3])
def test_group_replay_bitexact_hex(oracle, c3, sampler, smoother, group, kernel="wave"):
    m, tex = c3
    gm = GpuMesh.from_cube(m)
    _crew(gm, sampler, tex)
    om = oracle.OracleMesh.from_cube(m)
    _crew(om, sampler, tex)
    assert np.array_equal(om.permutation(), gm.permutation())
    if group > 0:
        gm.smooth_groups(smoother, 0, group, "fp64", kernel)
    before = gm.positions()
    counts = gm.smooth_groups(smoother, group, group + 1, "fp64", kernel)
    after = gm.positions()
    off, nodes = gm.groups()
    members = nodes[off[group]:off[group + 1]]
    rng = np.random.default_rng(300 + group)
    sample = np.sort(rng.choice(members, size=min(N_SAMPLE, members.size), replace=False)).astype(np.uint32)
    om.set_positions(before)
    ref_counts = om.smooth_vertices(smoother, sample)
    ref = om.positions()
    assert np.array_equal(after[sample], ref[sample])
    assert np.array_equal(counts[sample], ref_counts.astype(np.int32))
    outside = np.ones(m.nverts, bool)
    outside[members] = False
    assert np.array_equal(after[outside], before[outside])


@pytest.mark.parametrize("group", [0, 3])
def test_group_replay_plain_kernel_texture(oracle, c3, group):
    """The one-lane-per-vertex kernel under the 64^3 texture at full size
    (its patch walk is inlined into the kernel; out of line, its long
    branches once overwrote the function's return address)."""
    test_group_replay_bitexact_hex(oracle, c3, "texture", "nm", group, kernel="plain")


@pytest.mark.parametrize("sampler", ["analytic", "texture"])
def test_anisotropic_workload_run(c3, sampler):
    m, tex = c3
    runs = {}
    for precision in ("fp64", "fp32_polish"):
        gm = GpuMesh.from_cube(m)
        _crew(gm, sampler, tex)
        runs[precision] = gm.smooth("nm", 15, precision=precision, kernel="wave")
    r64, r32 = runs["fp64"], runs["fp32_polish"]
    for r in (r64, r32):
        assert r["min"][-1] > 0.0                      # no element inverted
        assert r["mean"][-1] > r["mean"][0] + 0.01      # the shell metric is followed
        assert np.all(np.diff(r["mean"]) > -1e-6)       # NM never lowers a patch's quality much
    assert abs(r32["mean"][-1] - r64["mean"][-1]) < 2e-4
    assert abs(r32["min"][-1] - r64["min"][-1]) < 5e-2
