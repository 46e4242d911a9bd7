"""The C-ABI library loads and exports every symbol include/gpumesh.h
declares, and the host-side entry points work without a GPU."""
import os
import re

import numpy as np
import pytest

from gpumesh_amd import _lib, cube_mesh
from gpumesh_amd.mesh import GpuMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gpumesh.h")).read()
    return sorted(set(re.findall(r"\b(gm_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert sorted(names) == sorted(_lib.EXPORTS)


def test_version_and_error_strings():
    L = _lib.lib()
    assert b"gfx950" in L.gm_version()
    n = _lib.device_count()
    assert n >= 0


def test_host_mesh_rejects_device_calls():
    m = cube_mesh(3, "tet", 0.2, seed=1)
    gm = GpuMesh.from_cube(m, device=False)
    with pytest.raises(_lib.GpuMeshError) as e:
        gm.evaluate()
    assert e.value.status == 1
    with pytest.raises(_lib.GpuMeshError) as e:
        gm.smooth("nm", 1)
    assert e.value.status == 1
    with pytest.raises(_lib.GpuMeshError, match="texture"):
        gm.set_crew("metric_conformity", "texture", 2.0, texture=([1, 1, 1], [0, 0, 0], [1, 1, 1], np.ones(6)))
    # Positions of a host mesh are its input positions in the new numbering.
    perm = gm.permutation()
    want = np.empty_like(m.pos)
    want[perm] = m.pos
    assert np.array_equal(gm.positions(), want)


def test_invalid_inputs_are_rejected():
    m = cube_mesh(2, "tet", 0.0, seed=1)
    bad = m.tets.copy()
    bad[0, 0] = m.nverts + 5
    with pytest.raises(_lib.GpuMeshError) as e:
        GpuMesh(m.pos, m.fixed, bad, m.hexs, device=False)
    assert e.value.status == 1
    with pytest.raises(_lib.GpuMeshError):
        GpuMesh(m.pos, m.fixed, m.tets.astype(np.int64) * 0 - 1, m.hexs, device=False)


def test_empty_mesh_builds():
    gm = GpuMesh(np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros((0, 4), np.uint32),
                 np.zeros((0, 8), np.uint32), device=False)
    assert gm.nverts == 0 and gm.ngroups == 0


def test_server_needs_a_device():
    from gpumesh_amd import GpuServer, device_count
    if device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(_lib.GpuMeshError) as e:
        GpuServer()
    assert e.value.status == 2


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the HIP library the product refuses to run (no CPU fallback):
    every entry point raises ImportError naming the missing file."""
    m = cube_mesh(2, "tet", 0.0, seed=1)
    missing = str(tmp_path / "libgpumesh_hip.so")
    monkeypatch.setattr(_lib, "LIB_PATH", missing)
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(ImportError, match="is not built"):
        GpuMesh.from_cube(m, device=False)
    with pytest.raises(ImportError, match="is not built"):
        _lib.math_probe(0, np.ones(4))
    with pytest.raises(ImportError, match="is not built"):
        cube_mesh(2, "tet", 0.0, seed=1)


def test_job_size_is_exact_beyond_two_gib():
    L = _lib.lib()
    small = int(L.gm_job_size(10, 4, 2))
    big = int(L.gm_job_size(2 ** 28, 2 ** 27, 0))
    # Sections grow linearly in the counts: positions 24 B, fixed flags 1 B
    # (8-byte aligned), tetrahedra 16 B per element.
    assert big > 2 ** 32
    grown = int(L.gm_job_size(2 ** 28 + 8, 2 ** 27 + 1, 0))
    assert grown - big == 8 * 24 + 8 + 16
    assert small > 0
