"""The device-side quality and optimiser code (gm_quality.h, gm_smoothers.h)
built for the host with g++ under the address and undefined-behaviour
sanitizers, on the cases the library refuses on the GPU: the texture sampler
with the one-lane-per-vertex Nelder-Mead functor (metric conformity on
hexahedra) and the GETMe hexahedron element step (metric-wise volumes through
the adaptive walk).  The host build must agree bit for bit with the oracle,
raise no sanitizer report, and every walk must end within a few thousand
iterations: the arithmetic of those paths terminates and stays in bounds."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpumesh_amd import _lib, cube_mesh
from gpumesh_amd.mesh import GpuMesh
from gpumesh_amd.samplers import shell_texture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "device_path.cpp")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("devpath") / "device_path")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                    "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all",
                    "-o", out, SRC], check=True)
    return out


@pytest.mark.parametrize("n", [4, 5, 7])
def test_texture_paths_on_host(oracle, exe, tmp_path, n):
    m = cube_mesh(n, "hex", 0.3, seed=3)
    K = float(n)
    size, tmin, text, tex = shell_texture(m.pos, 16, K)
    gm = GpuMesh.from_cube(m, device=False)
    L = _lib.lib()
    eoff = np.zeros(gm.nverts + 1, np.uint32)
    eref = np.zeros(gm.sum_patch, np.uint32)
    noff = np.zeros(gm.nverts + 1, np.uint32)
    nbr = np.zeros(gm.sum_nbr, np.uint32)
    _lib.check(L.gm_mesh_topology(gm._h, _lib.ptr(eoff), _lib.ptr(eref), _lib.ptr(noff), _lib.ptr(nbr)), "topology")
    hexs = np.zeros(8 * gm.nhexs, np.uint32)
    _lib.check(L.gm_mesh_elements(gm._h, None, _lib.ptr(hexs)), "elements")
    _, nodes = gm.groups()
    perm = gm.permutation()
    pos = np.empty_like(m.pos)
    pos[perm] = m.pos
    d = str(tmp_path)
    for name, a in [("pos", pos), ("hexs", hexs), ("eoff", eoff), ("eref", eref), ("noff", noff), ("nbr", nbr),
                    ("order", nodes[:gm.nmovable].astype(np.uint32)), ("tsize", size.astype(np.int32)),
                    ("box", np.r_[tmin, text, K]), ("texels", tex.astype(np.float32))]:
        np.ascontiguousarray(a).tofile(os.path.join(d, name + ".bin"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, d], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "runtime error" not in r.stderr, r.stderr[-2000:]

    om = oracle.OracleMesh.from_cube(m)
    om.set_crew("metric_conformity", "texture", scaling=K, texture=(size, tmin, text, tex))
    assert np.array_equal(om.permutation(), perm)
    # Element step against the oracle's transform of the same hexahedra.
    vpp = np.fromfile(os.path.join(d, "getme_vpp.bin"), np.float64).reshape(-1, 8, 3)
    w = np.fromfile(os.path.join(d, "getme_w.bin"), np.float64)
    iters = np.fromfile(os.path.join(d, "walk_iters.bin"), np.int64)
    for e in range(gm.nhexs):
        ov, ow = om.getme_element(pos[hexs[8 * e: 8 * e + 8]])
        assert np.array_equal(vpp[e], ov) and w[e] == ow, e
    assert iters.max() < 20000 and iters.min() > 0
    # Nelder-Mead passes against the oracle's.
    ref = om.smooth("nm", 2, threads=1)
    nm = np.fromfile(os.path.join(d, "nm_pos.bin"), np.float64).reshape(-1, 3)
    counts = np.fromfile(os.path.join(d, "nm_counts.bin"), np.int32)
    assert np.array_equal(nm, om.positions())
    assert np.array_equal(counts, ref["eval_counts"].astype(np.int32))
