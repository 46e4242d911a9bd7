"""FP64 GPU path vs the CPU oracle on the same seeded meshes.

The FP64 kernels are built without contraction and use the same operation
order and the same transcendentals (gm_libm) as the oracle in libm mode 1, so
positions, qualities and per-vertex patch-evaluation counts must be
bit-identical; the min/mean reductions use a different (compensated tree)
summation order, so the harmonic mean is compared within 1e-12.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh, math_probe

pytestmark = pytest.mark.gpu

MEAN_TOL = 1e-12


def _pair(oracle, n, kind, jitter, seed):
    m = cube_mesh(n, kind, jitter, seed=seed)
    return m, oracle.OracleMesh.from_cube(m), GpuMesh.from_cube(m)


def test_math_probe_bitexact(oracle):
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.uniform(0, 10, 4000), 10.0 ** rng.uniform(-300, 300, 4000),
                        np.array([0.0, 1.0, 8.0, 27.0, 5e-324, 2.2250738585072014e-308, 1e308])])
    y = rng.uniform(0.1, 10, x.size)
    assert np.array_equal(math_probe(0, x).view(np.uint64), np.sqrt(x).view(np.uint64))
    assert np.array_equal(math_probe(1, x, y).view(np.uint64), (x / y).view(np.uint64))
    assert np.array_equal(math_probe(2, x).view(np.uint64), oracle.gm_libm(2, x).view(np.uint64))
    assert np.array_equal(math_probe(3, x).view(np.uint64), oracle.gm_libm(3, x).view(np.uint64))
    e = rng.uniform(-700, 700, 4000)
    assert np.array_equal(math_probe(4, e).view(np.uint64), oracle.gm_libm(4, e).view(np.uint64))


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_element_qualities_bitexact(oracle, kind):
    m, om, gm = _pair(oracle, 8, kind, 0.3, 11)
    g = gm.evaluate(element_qualities=True)
    r = om.evaluate()
    assert np.array_equal(g["elem_qual"], om.element_qualities())
    assert g["min"] == r["min"]
    assert abs(g["mean"] - r["mean"]) <= MEAN_TOL
    assert np.array_equal(g["hist"], r["hist"])


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kernel", ["plain", "wave"])
@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_smoothing_fp64_bitexact(oracle, smoother, kernel, kind):
    if smoother == "getme" and kernel == "wave":
        pytest.skip("GETMe has a single kernel shape")
    n = 10 if kind == "tet" else 12
    passes = 3
    m, om, gm = _pair(oracle, n, kind, 0.3, 5)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    ref = om.smooth(smoother, passes, threads=4, spawn_coeff=coeff)
    out = gm.smooth(smoother, passes, precision="fp64", kernel=kernel, spawn_coeff=coeff)
    p_ref, p_gpu = om.positions(), gm.positions()
    assert np.array_equal(p_gpu, p_ref), f"max |d| = {np.max(np.abs(p_gpu - p_ref))}"
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])
    assert np.max(np.abs(out["mean"] - ref["mean"])) <= MEAN_TOL
    assert np.array_equal(out["hist"], ref["hist"])
