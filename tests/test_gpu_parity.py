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
    c = np.concatenate([rng.uniform(-40, 40, 4000), rng.uniform(-1e5, 1e5, 500)])
    assert np.array_equal(math_probe(6, c).view(np.uint64), oracle.gm_libm(6, c).view(np.uint64))
    assert np.array_equal(math_probe(7, x).view(np.uint64), oracle.gm_libm(7, x).view(np.uint64))
    assert np.array_equal(math_probe(8, y, x[:y.size] % 7).view(np.uint64),
                          oracle.gm_libm(8, y, x[:y.size] % 7).view(np.uint64))


def test_math_probe_fast_powm23_bound():
    """a^(-2/3) of the mixed-precision kernels (FP32 hardware seed, one
    Halley step): within 4e-15 relative of an 80-bit reference over the float
    range, 1e-15 in the unit range; the exact routine beyond the float range
    and at 0 (inf), inf (0) and negative input (NaN)."""
    rng = np.random.default_rng(3)
    x = np.concatenate([2.0 ** rng.uniform(-126, 127, 20000), rng.uniform(1e-9, 2.0, 20000)])
    got = math_probe(5, x)
    ref = np.power(x.astype(np.longdouble), np.longdouble(-2) / np.longdouble(3))
    rel = np.abs((got.astype(np.longdouble) - ref) / ref).astype(np.float64)
    assert rel.max() < 4e-15, rel.max()
    unit = (x > 1e-9) & (x < 2.0)
    assert rel[unit].max() < 1e-15, rel[unit].max()
    edge = np.array([0.0, np.inf, -1.0, 1e-300, 1e300])
    g = math_probe(5, edge)
    assert g[0] == np.inf and g[1] == 0.0 and np.isnan(g[2])
    ref = np.power(edge[3:].astype(np.longdouble), np.longdouble(-2) / np.longdouble(3))
    assert np.all(np.abs((g[3:].astype(np.longdouble) - ref) / ref) < 1e-15)


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


def _crew_pair(oracle, kind, n, evaluator, sampler, K, seed=3, texture_n=16):
    from gpumesh_amd.samplers import shell_texture
    m, om, gm = _pair(oracle, n, kind, 0.3, seed)
    tex = shell_texture(m.pos, texture_n, K) if sampler == "texture" else None
    om.set_crew(evaluator, sampler, scaling=K, texture=tex)
    gm.set_crew(evaluator, sampler, scaling=K, texture=tex)
    if sampler == "analytic_reference":
        om.set_aspect_ratio(4.0)
        gm.set_aspect_ratio(4.0)
    return m, om, gm


@pytest.mark.parametrize("sampler", ["analytic", "texture", "analytic_reference"])
@pytest.mark.parametrize("evaluator", ["metric_conformity", "mean_ratio"])
def test_anisotropic_element_qualities_bitexact(oracle, evaluator, sampler):
    """Metric-conformity and metric-wise mean-ratio qualities under the shell
    field, analytic and sampled (FP32 texels, FP64 trilinear)."""
    for kind in ("hex", "tet"):
        m, om, gm = _crew_pair(oracle, kind, 6, evaluator, sampler, K=6.0)
        g = gm.evaluate(element_qualities=True)
        assert np.array_equal(g["elem_qual"], om.element_qualities())
        r = om.evaluate()
        assert g["min"] == r["min"] and abs(g["mean"] - r["mean"]) <= MEAN_TOL


@pytest.mark.parametrize("smoother,kernel", [("nm", "wave"), ("nm", "plain"), ("gd", "wave"), ("spawn", "wave"),
                                             ("getme", "plain"), ("getme", "wave")])
@pytest.mark.parametrize("sampler", ["analytic", "texture", "analytic_reference"])
def test_anisotropic_smoothing_fp64_bitexact(oracle, smoother, kernel, sampler):
    """The anisotropic hex workload's crew (metric conformity, metric-wise
    measurer) on a small jittered hex cube; every kernel shape, every sampler."""
    m, om, gm = _crew_pair(oracle, "hex", 4, "metric_conformity", sampler, K=4.0)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    ref = om.smooth(smoother, 2, threads=4, spawn_coeff=coeff)
    out = gm.smooth(smoother, 2, precision="fp64", kernel=kernel, spawn_coeff=coeff)
    p_ref, p_gpu = om.positions(), gm.positions()
    assert np.array_equal(p_gpu, p_ref), f"max |d| = {np.max(np.abs(p_gpu - p_ref))}"
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])


@pytest.mark.parametrize("sampler", ["analytic", "analytic_reference"])
def test_metric_wise_mean_ratio_smoothing_bitexact(oracle, sampler):
    """Mean ratio through the metric-wise measurer (adaptive Riemannian
    segments) under an analytic field, GD and GETMe on a Kuhn cube."""
    for smoother in ("gd", "getme"):
        m, om, gm = _crew_pair(oracle, "tet", 4, "mean_ratio", sampler, K=4.0)
        om.smooth(smoother, 2, threads=4)
        gm.smooth(smoother, 2, precision="fp64", kernel="wave")
        assert np.array_equal(gm.positions(), om.positions())


@pytest.mark.parametrize("smoother,kernel", [("nm", "plain"), ("getme", "plain"), ("getme", "wave")])
def test_texture_sampler_larger_hex_cubes(oracle, smoother, kernel):
    """The 7^3 (NM, plain kernel) and 5^3 (GETMe) hex cubes under metric
    conformity with the texture sampler, the runs whose kernels never ended
    while the plain patch functor was kept out of line (its long branches
    went through its return-address registers)."""
    n = 7 if smoother == "nm" else 5
    m, om, gm = _crew_pair(oracle, "hex", n, "metric_conformity", "texture", K=float(n))
    ref = om.smooth(smoother, 2, threads=4)
    out = gm.smooth(smoother, 2, precision="fp64", kernel=kernel)
    assert np.array_equal(gm.positions(), om.positions())
    assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(out["min"], ref["min"])


def test_mixed_element_mesh_bitexact(oracle):
    """A mesh holding tets and hexes uses the generic kernels (patches mix
    element kinds); positions must still match."""
    t = cube_mesh(4, "tet", 0.0, seed=1)
    h = cube_mesh(4, "hex", 0.0, seed=1)
    # Right half of the cube from tets, left half from hexes, sharing the grid.
    cx_t = t.pos[t.tets].mean(axis=1)[:, 0]
    cx_h = h.pos[h.hexs].mean(axis=1)[:, 0]
    tets = t.tets[cx_t > 0.5]
    hexs = h.hexs[cx_h < 0.5]
    rng = np.random.default_rng(2)
    pos = t.pos + np.where(t.fixed[:, None] == 1, 0.0, rng.uniform(-0.05, 0.05, t.pos.shape))
    om = oracle.OracleMesh(pos, t.fixed, tets, hexs)
    gm = GpuMesh(pos, t.fixed, tets, hexs)
    for smoother in ("nm", "gd", "spawn", "getme"):
        om.set_positions(om.positions())
        ref = om.smooth(smoother, 2, threads=2)
        out = gm.smooth(smoother, 2, precision="fp64", kernel="wave")
        assert np.array_equal(gm.positions(), om.positions()), smoother
        assert np.array_equal(out["eval_counts"], ref["eval_counts"].astype(np.int32))


@pytest.mark.parametrize("kernel", ["plain", "wave"])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_c1_every_smoother_against_fixture(smoother, kernel):
    """Every smoother on the 20^3 plain-method mesh in FP64: the committed
    fixture's positions after passes 1, 2 and 10, minima, histograms and
    evaluation count (tests/golden/c1_smoothers.json)."""
    import hashlib
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                    "c1_smoothers.json")))["smoothers"][smoother]
    gm = GpuMesh.from_cube(cube_mesh(20, "tet", 0.3, seed=1))
    mins, evals = [], 0
    for upto, step in ((1, 1), (2, 1), (10, 8)):
        out = gm.smooth(smoother, step, precision="fp64", kernel=kernel)
        mins += out["min"].tolist() if not mins else out["min"][1:].tolist()
        evals += int(out["eval_counts"].sum())
        pos = gm.positions()
        assert hashlib.sha256(np.ascontiguousarray(pos).tobytes()).hexdigest() == g["positions_sha256"][str(upto)]
        assert out["hist"][-1].tolist() == g["hist"][str(upto)]
    assert mins == g["min"] and evals == g["eval_count_total"]


def test_c1_plain_method_workload_against_fixture(oracle):
    """The plain-method workload (20^3 Kuhn cube, jitter 0.3 h, seed 1, NM x 10,
    FP64) on the GPU reproduces the committed oracle fixture bit for bit."""
    import hashlib
    import json
    import os
    golden = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_golden.json")))
    m = cube_mesh(20, "tet", 0.3, seed=1)
    for kernel in ("plain", "wave"):
        gm = GpuMesh.from_cube(m)
        out = gm.smooth("nm", 10, precision="fp64", kernel=kernel)
        pos = gm.positions()
        assert hashlib.sha256(np.ascontiguousarray(pos).tobytes()).hexdigest() == \
            golden["gm_libm"]["final_positions_sha256"], kernel
        assert out["min"].tolist() == golden["gm_libm"]["min"]
        assert np.max(np.abs(out["mean"] - np.array(golden["gm_libm"]["mean"]))) <= MEAN_TOL
        assert int(out["eval_counts"].sum()) == golden["gm_libm"]["eval_count_total"]
