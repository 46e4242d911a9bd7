"""The reference-side binding on the GPU: include/gpumesh_binding.hpp
(compiled into tests/host/_build/binding_test by __graft_entry__.build())
runs the reference's loop -- an initial evaluation, then per pass a record
and one sweep -- on a mesh as NodeGroups clustered it, and must give the
oracle's positions bit for bit with every position on its own vertex; and
a mesh in any numbering with the reference's groups mapped onto it
(gm_mesh_create_grouped) smooths exactly as the renumbering entry does."""
import os
import subprocess

import numpy as np
import pytest

from gpumesh_amd import cube_mesh
from gpumesh_amd.mesh import GpuMesh

from test_binding import EXE, clustered, clustered_box, write_inputs

pytestmark = pytest.mark.gpu

SMOOTHERS = {"nm": 0, "gd": 1, "spawn": 2, "getme": 3}


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kind,n", [("tet", 8), ("hex", 6)])
def test_binding_run_matches_oracle(oracle, tmp_path, kind, n, smoother):
    assert os.path.exists(EXE), "tests/host/_build/binding_test is built by __graft_entry__.build()"
    m, om, pos, fixed, tets, hexs, goff, gnodes = clustered(oracle, kind, n)
    d = str(tmp_path)
    write_inputs(d, pos, fixed, tets, hexs, goff, gnodes)
    passes = 3
    r = subprocess.run([EXE, "device", d, str(SMOOTHERS[smoother]), str(passes), "0"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = om.smooth(smoother, passes, threads=4)
    got = np.fromfile(os.path.join(d, "out_pos.bin"), np.float64).reshape(-1, 3)
    assert np.array_equal(got, om.positions())
    stats = np.fromfile(os.path.join(d, "out_stats.bin"), np.float64).reshape(-1, 3)
    hist = np.fromfile(os.path.join(d, "out_hist.bin"), np.int32).reshape(-1, 20)
    assert stats.shape[0] == passes + 1
    assert np.array_equal(stats[:, 0], ref["min"])
    assert np.allclose(stats[:, 1], ref["inv_sum"], rtol=1e-12, atol=0)
    assert np.array_equal(hist, ref["hist"])


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
@pytest.mark.parametrize("kind,n", [("tet", 7), ("hex", 6)])
def test_binding_sliding_box_matches_oracle(oracle, tmp_path, kind, n, smoother):
    """The binding on a BoxBoundary mesh whose faces and edges slide, with
    the constraints described by their geometry: the oracle's positions and
    pass records."""
    assert os.path.exists(EXE), "tests/host/_build/binding_test is built by __graft_entry__.build()"
    m, om, pos, fixed, tets, hexs, goff, gnodes, cons, ctab = clustered_box(oracle, kind, n)
    d = str(tmp_path)
    write_inputs(d, pos, fixed, tets, hexs, goff, gnodes, cons, ctab)
    r = subprocess.run([EXE, "device", d, str(SMOOTHERS[smoother]), "2", "0"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    ref = om.smooth(smoother, 2, threads=4)
    got = np.fromfile(os.path.join(d, "out_pos.bin"), np.float64).reshape(-1, 3)
    assert np.array_equal(got, om.positions())
    stats = np.fromfile(os.path.join(d, "out_stats.bin"), np.float64).reshape(-1, 3)
    assert np.array_equal(stats[:, 0], ref["min"])
    assert np.array_equal(np.fromfile(os.path.join(d, "out_hist.bin"), np.int32).reshape(-1, 20), ref["hist"])


@pytest.mark.parametrize("smoother", ["nm", "gd", "spawn", "getme"])
def test_binding_global_pass_matches_oracle(oracle, tmp_path, smoother):
    """One global pass of the reference's loop through the binding: BATR on
    the device handle (Session::restructure), the reference's renumbering,
    then FP64 sweeps: the oracle's restructured mesh and positions."""
    assert os.path.exists(EXE), "tests/host/_build/binding_test is built by __graft_entry__.build()"
    m, om, pos, fixed, tets, hexs, goff, gnodes = clustered(oracle, "tet", 7)
    d = str(tmp_path)
    write_inputs(d, pos, fixed, tets, hexs, goff, gnodes)
    r = subprocess.run([EXE, "global", d, "7.0", str(SMOOTHERS[smoother]), "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    om.set_crew("mean_ratio", "uniform", scaling=7.0)
    want = om.restructure(5, 5, 0.5, 1.5)
    counts = np.fromfile(os.path.join(d, "out_counts.bin"), np.int64)
    assert list(counts) == [want[k] for k in ("face_swaps", "edge_swaps", "splits", "merges", "passes")]
    om.smooth(smoother, 2, threads=4)
    assert np.array_equal(np.fromfile(os.path.join(d, "out_tets.bin"), np.uint32).reshape(-1, 4), om.elements()[0])
    assert np.array_equal(np.fromfile(os.path.join(d, "out_pos.bin"), np.float64).reshape(-1, 3), om.positions())


@pytest.mark.parametrize("kind,n", [("tet", 9), ("hex", 7)])
def test_grouped_mesh_in_caller_numbering(oracle, kind, n):
    """The original numbering, the reference's groups mapped to it, each
    group's members shuffled: FP64 sweeps equal the oracle's, mapped."""
    m = cube_mesh(n, kind, 0.3, seed=6)
    om = oracle.OracleMesh.from_cube(m)
    perm = om.permutation()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(m.nverts)
    groups, _ = om.groups()
    rng = np.random.default_rng(1)
    nodes = np.concatenate([rng.permutation(inv[np.arange(g[4], g[5])]) for g in groups]).astype(np.uint32)
    off = np.concatenate([[0], np.cumsum([g[5] - g[4] for g in groups])]).astype(np.int64)
    gm = GpuMesh.grouped(m.pos, m.fixed, m.tets, m.hexs, off, nodes)
    for smoother in ("nm", "gd", "spawn", "getme"):
        start = gm.positions()
        om.set_positions(start[inv])
        ref = om.smooth(smoother, 2, threads=4)
        out = gm.smooth(smoother, 2, precision="fp64", kernel="wave")
        assert np.array_equal(gm.positions(), om.positions()[perm]), smoother
        assert np.array_equal(out["eval_counts"], ref["eval_counts"][perm].astype(np.int32)), smoother
        assert np.array_equal(out["min"], ref["min"]), smoother
