"""The autopilot schedule on the device (gm_mesh_smooth_schedule,
gm_batch_smooth_schedule) against the oracle's restatement of
AbstractSmoother::evaluateMeshQuality (Smoothers/AbstractSmoother.cpp:186-197):
the stop is decided on the device after each pass's quality reduction, so in
FP64 the number of sweeps, the minima, histograms, evaluation counts and the
final positions must be the oracle's bits (the harmonic means differ only by
the rounding of the reduction order)."""
import numpy as np
import pytest

from gpumesh_amd import GpuBatch, GpuMesh, cube_mesh
from gpumesh_amd._lib import GpuMeshError

pytestmark = pytest.mark.gpu

T_MEAN, T_MIN = 2e-4, 5e-3     # Schedule::minQualThreshold, Schedule::qualMeanThreshold


@pytest.mark.parametrize("kernel", ["wave", "plain"])
@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_autopilot_fp64_bitexact(oracle, smoother, kernel):
    m = cube_mesh(10, "tet", 0.35, seed=4)
    om = oracle.OracleMesh.from_cube(m)
    ref = om.smooth_autopilot(smoother, 15, min_qual_threshold=T_MEAN, qual_mean_threshold=T_MIN, threads=8)
    gm = GpuMesh.from_cube(m)
    r = gm.smooth_autopilot(smoother, 15, min_qual_threshold=T_MEAN, qual_mean_threshold=T_MIN, precision="fp64",
                            kernel=kernel)
    print(f"{smoother}/{kernel}: {r['passes']} sweeps (oracle {ref['passes']})", flush=True)
    assert r["passes"] == ref["passes"]
    assert 1 <= r["passes"] < 15
    assert np.array_equal(r["min"], ref["min"])
    # The device's compensated tree sum of 1/q and the oracle's serial sum
    # round differently (the harmonic means agree to about 1e-16).
    assert np.allclose(r["mean"], ref["mean"], rtol=0.0, atol=1e-12)
    assert np.array_equal(r["hist"], ref["hist"])
    assert np.array_equal(r["eval_counts"], ref["eval_counts"].astype(np.int32))
    assert np.array_equal(gm.positions(), om.positions())


def test_autopilot_cap_and_zero(oracle):
    m = cube_mesh(6, "hex", 0.3, seed=5)
    gm = GpuMesh.from_cube(m)
    assert gm.smooth_autopilot("gd", 0)["passes"] == 0
    gm = GpuMesh.from_cube(m)
    r = gm.smooth_autopilot("gd", 4, min_qual_threshold=-1.0)      # never stops by itself
    assert r["passes"] == 4 and r["min"].size == 5


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autopilot_batch_members_match_alone(precision, graph):
    specs = [(8, "tet", 0.05, 11), (6, "hex", 0.3, 12), (12, "tet", 0.35, 13), (10, "tet", 0.45, 14)]
    meshes = [cube_mesh(*s) for s in specs]
    alone = []
    for m in meshes:
        g = GpuMesh.from_cube(m)
        alone.append((g.smooth_autopilot("nm", 12, T_MEAN, T_MIN, precision=precision), g.positions()))
    gms = [GpuMesh.from_cube(m) for m in meshes]
    b = GpuBatch(gms)
    r = b.smooth_autopilot("nm", 12, T_MEAN, T_MIN, precision=precision, graph=graph)
    for i, (a, pos) in enumerate(alone):
        k = a["passes"]
        assert r["passes"][i] == k
        assert np.array_equal(r["min"][i, :k + 1], a["min"])
        assert np.array_equal(r["mean"][i, :k + 1], a["mean"])
        assert np.array_equal(gms[i].positions(), pos)
    assert len(set(int(p) for p in r["passes"])) > 1, "members should stop at different passes"
    b.close()


def test_autopilot_rejects_polish():
    gm = GpuMesh.from_cube(cube_mesh(4, "tet", 0.3, seed=1))
    with pytest.raises(GpuMeshError):
        gm.smooth_autopilot("gd", 5, precision="fp32_polish")
    # A fixed-count run afterwards is unaffected by the rejected call.
    r = gm.smooth("gd", 2, precision="fp64", kernel="wave")
    assert r["min"].size == 3
