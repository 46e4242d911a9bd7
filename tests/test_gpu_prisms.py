"""Prisms on the device against the oracle: meshes of prisms only and of
hexahedra, prisms and tetrahedra together (MeshPri quality, volumes and GETMe
transform, AbstractEvaluator::patchQuality's prism case).  FP64 runs are
bit-identical to the oracle (positions, evaluation counts, minimum); the
harmonic mean agrees within 1e-12 (reduction order).  The mixed-precision
modes are checked against the FP64 oracle within stated bounds, and batches
and the server against the same meshes alone."""
import numpy as np
import pytest

from gpumesh_amd import GpuBatch, GpuMesh, GpuServer
from gpumesh_amd.meshgen import element_cube_mesh

pytestmark = pytest.mark.gpu

MEAN_TOL = 1e-12
COEFF = {"nm": 1.0 / 20.0, "gd": 1.0 / 20.0, "spawn": 0.25, "getme": 1.0 / 20.0}


def _pair(oracle, layout, n, seed, jitter=0.3):
    m = element_cube_mesh(n, layout, jitter, seed)
    return m, oracle.OracleMesh.from_cube(m), GpuMesh.from_cube(m)


@pytest.mark.parametrize("layout", ["prism", "mixed"])
@pytest.mark.parametrize("evaluator,sampler", [("mean_ratio", "uniform"), ("metric_conformity", "uniform"),
                                               ("metric_conformity", "analytic"), ("mean_ratio", "analytic")])
def test_prism_element_qualities_bitexact(oracle, layout, evaluator, sampler):
    m, om, gm = _pair(oracle, layout, 5, 2)
    om.set_crew(evaluator, sampler, scaling=5.0)
    gm.set_crew(evaluator, sampler, scaling=5.0)
    g = gm.evaluate(element_qualities=True)
    assert np.array_equal(g["elem_qual"], om.element_qualities())
    r = om.evaluate()
    assert g["min"] == r["min"] and abs(g["mean"] - r["mean"]) <= MEAN_TOL
    assert np.array_equal(g["hist"], r["hist"])


@pytest.mark.parametrize("kernel", ["plain", "wave"])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("layout", ["prism", "mixed"])
def test_prism_smoothing_fp64_bitexact(oracle, layout, smoother, kernel):
    m, om, gm = _pair(oracle, layout, 5, 3)
    ref = om.smooth(smoother, 3, threads=4, spawn_coeff=COEFF[smoother])
    out = gm.smooth(smoother, 3, precision="fp64", kernel=kernel, spawn_coeff=COEFF[smoother])
    p_ref, p_gpu = om.positions(), gm.positions()
    assert np.array_equal(p_gpu, p_ref), f"max |d| = {np.max(np.abs(p_gpu - p_ref))}"
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])
    assert np.max(np.abs(out["mean"] - ref["mean"])) <= MEAN_TOL


@pytest.mark.parametrize("smoother", ["nm", "gd", "getme"])
def test_prism_metric_wise_smoothing_bitexact(oracle, smoother):
    """Metric conformity under the analytic shell field (metric-wise
    measurer: adaptive Riemannian segments in the prism volumes of GETMe)."""
    m, om, gm = _pair(oracle, "mixed", 4, 6)
    om.set_crew("metric_conformity", "analytic", scaling=4.0)
    gm.set_crew("metric_conformity", "analytic", scaling=4.0)
    om.smooth(smoother, 2, threads=4)
    gm.smooth(smoother, 2, precision="fp64", kernel="wave")
    assert np.array_equal(gm.positions(), om.positions())


@pytest.mark.parametrize("precision", ["fp32", "fp32_polish", "fp64_fast"])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_prism_mixed_precision_against_fp64_oracle(oracle, smoother, precision):
    """FP32 passes (generic kernels on prisms) against the FP64 oracle:
    final mean within 2e-3, minimum within 2e-2, positions within h/2."""
    m, om, gm = _pair(oracle, "mixed", 6, 4)
    ref = om.smooth(smoother, 4, threads=4, spawn_coeff=COEFF[smoother])
    out = gm.smooth(smoother, 4, precision=precision, kernel="wave", spawn_coeff=COEFF[smoother])
    assert abs(out["mean"][-1] - ref["mean"][-1]) < 2e-3
    assert abs(out["min"][-1] - ref["min"][-1]) < 2e-2
    assert np.max(np.abs(gm.positions() - om.positions())) < 0.5 / 6
    assert out["mean"][-1] > out["mean"][0]


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("smoother", ["nm", "spawn", "getme"])
def test_prism_batch_members_match_alone(smoother, graph):
    meshes = [element_cube_mesh(n, lay, 0.3, s) for n, lay, s in ((4, "prism", 1), (6, "mixed", 2), (3, "prism", 3))]
    alone = []
    for m in meshes:
        g = GpuMesh.from_cube(m)
        r = g.smooth(smoother, 2, precision="fp32_polish", kernel="wave", spawn_coeff=COEFF[smoother])
        alone.append((g.positions(), r))
    ms = [GpuMesh.from_cube(m) for m in meshes]
    b = GpuBatch(ms)
    r = b.smooth(smoother, 2, precision="fp32_polish", kernel="wave", spawn_coeff=COEFF[smoother], graph=graph)
    for i, g in enumerate(ms):
        assert np.array_equal(g.positions().view(np.uint64), alone[i][0].view(np.uint64)), i
        assert np.array_equal(r["inv_sum"][i], alone[i][1]["inv_sum"]), i
    b.close()
