"""Meshes that start with inverted elements: FP64 GPU path vs the oracle.

Three interior vertices of a jittered 6^3 cube are pushed past their
neighbours (about 1.3 h), which inverts 11-14 elements.  Patches then hold
non-positive corners, so the smoothers take the exact rule (the patch
quality sticks to the most negative element) instead of the harmonic mean,
the candidate sweeps compare inverted candidates, and the histograms see
negative qualities.  The oracle untangles every case within three passes.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh

pytestmark = pytest.mark.gpu

MEAN_TOL = 1e-12


def tangled_mesh(kind):
    m = cube_mesh(6, kind, 0.2, seed=4)
    pos = m.pos.copy()
    h = 1.0 / 6
    interior = np.where(m.fixed == 0)[0]
    pick = interior[[3, len(interior) // 2, len(interior) - 5]]
    pos[pick] += h * np.array([[1.3, 0.2, -0.1], [-0.2, 1.25, 0.3], [0.1, -0.2, -1.3]])
    return m, pos


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kernel", ["plain", "wave"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_tangled_fp64_bitexact(oracle, smoother, kernel, kind):
    m, pos = tangled_mesh(kind)
    om = oracle.OracleMesh(pos, m.fixed, m.tets, m.hexs)
    gm = GpuMesh(pos, m.fixed, m.tets, m.hexs)
    g0 = gm.evaluate(element_qualities=True)
    assert np.array_equal(g0["elem_qual"], om.element_qualities())
    assert g0["min"] < 0.0
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    ref = om.smooth(smoother, 3, threads=2, spawn_coeff=coeff)
    out = gm.smooth(smoother, 3, precision="fp64", kernel=kernel, spawn_coeff=coeff)
    p_ref, p_gpu = om.positions(), gm.positions()
    assert np.array_equal(p_gpu, p_ref), f"max |d| = {np.max(np.abs(p_gpu - p_ref))}"
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])
    assert np.max(np.abs(out["mean"] - ref["mean"])) <= MEAN_TOL
    assert np.array_equal(out["hist"], ref["hist"])
    assert out["min"][-1] > 0.0


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_tangled_mixed_precision_untangles(smoother, kind):
    """FP32 with the FP64 polish pass also untangles the mesh, and its final
    mean stays within 0.02 of the FP64 run's (the trajectories part in the
    first pass, where the moves are large)."""
    m, pos = tangled_mesh(kind)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    r64 = GpuMesh(pos, m.fixed, m.tets, m.hexs).smooth(smoother, 3, precision="fp64", kernel="wave",
                                                       spawn_coeff=coeff)
    r32 = GpuMesh(pos, m.fixed, m.tets, m.hexs).smooth(smoother, 3, precision="fp32_polish", kernel="wave",
                                                       spawn_coeff=coeff)
    # FP32 runs record the start from the FP32-rounded positions.
    assert abs(r32["min"][0] - r64["min"][0]) < 1e-6 and r32["min"][0] < 0.0
    assert r32["min"][-1] > 0.0, r32["min"]
    assert abs(r32["mean"][-1] - r64["mean"][-1]) < 0.02, (r32["mean"][-1], r64["mean"][-1])


@pytest.mark.parametrize("kind,evaluator,smoother,kernel", [
    ("hex", "metric_conformity", "nm", "wave"), ("hex", "metric_conformity", "nm", "plain"),
    ("hex", "metric_conformity", "gd", "wave"), ("hex", "metric_conformity", "spawn", "wave"),
    ("hex", "metric_conformity", "getme", "wave"),
    ("tet", "mean_ratio", "gd", "wave"), ("tet", "mean_ratio", "getme", "wave")])
def test_tangled_anisotropic_fp64_bitexact(oracle, kind, evaluator, smoother, kernel, sampler="analytic"):
    """The same tangles under the shell field (analytic sampler, metric-wise
    measurer): metric conformity signs inverted elements by det(F), the
    metric-wise mean ratio measures adaptive Riemannian segments (wave
    kernels only: the one-lane-per-vertex kernels walk those segments for a
    whole patch in one lane)."""
    from gpumesh_amd.samplers import shell_texture
    m, pos = tangled_mesh(kind)
    om = oracle.OracleMesh(pos, m.fixed, m.tets, m.hexs)
    gm = GpuMesh(pos, m.fixed, m.tets, m.hexs)
    tex = shell_texture(pos, 16, 4.0) if sampler == "texture" else None
    om.set_crew(evaluator, sampler, scaling=4.0, texture=tex)
    gm.set_crew(evaluator, sampler, scaling=4.0, texture=tex)
    g0 = gm.evaluate(element_qualities=True)
    assert np.array_equal(g0["elem_qual"], om.element_qualities())
    assert g0["min"] < 0.0
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    ref = om.smooth(smoother, 2, threads=2, spawn_coeff=coeff)
    out = gm.smooth(smoother, 2, precision="fp64", kernel=kernel, spawn_coeff=coeff)
    p_ref, p_gpu = om.positions(), gm.positions()
    assert np.array_equal(p_gpu, p_ref), f"max |d| = {np.max(np.abs(p_gpu - p_ref))}"
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])


@pytest.mark.parametrize("smoother,kernel", [("nm", "wave"), ("gd", "wave"), ("spawn", "wave"), ("getme", "wave"),
                                             ("nm", "plain"), ("getme", "plain")])
def test_tangled_texture_fp64_bitexact(oracle, smoother, kernel):
    """The tangled hex cube under metric conformity with the texture sampler
    (the shell field sampled at 16^3 cell centres in FP32), every kernel
    shape."""
    test_tangled_anisotropic_fp64_bitexact(oracle, "hex", "metric_conformity", smoother, kernel, sampler="texture")
