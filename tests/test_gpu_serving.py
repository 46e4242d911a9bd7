"""Serving loop: jobs that arrive while others run are admitted as memory
and slots free up and advance together in fused rounds; every result must
be bit-identical to smoothing that mesh alone with gm_mesh_smooth (same
smoother, passes and precision), whatever else was in the mix."""
import time

import numpy as np
import pytest

from gpumesh_amd import GpuMesh, GpuServer, cube_mesh
from gpumesh_amd._lib import GpuMeshError

pytestmark = pytest.mark.gpu

JOBS = [  # (n, kind, jitter, seed, smoother, passes, precision)
    (9, "tet", 0.3, 101, "nm", 4, "fp32_polish"),
    (14, "hex", 0.25, 102, "gd", 6, "fp32_polish"),
    (6, "tet", 0.4, 103, "getme", 5, "fp32"),
    (17, "tet", 0.2, 104, "gd", 3, "fp64"),
    (11, "hex", 0.35, 105, "getme", 7, "fp32_polish"),
    (4, "tet", 0.3, 106, "spawn", 2, "fp64"),
    (20, "tet", 0.3, 107, "getme", 5, "fp32_polish"),
    (8, "hex", 0.3, 108, "nm", 3, "fp32"),
    (9, "tet", 0.3, 109, "nm", 1, "fp32_polish"),
    (12, "tet", 0.3, 110, "gd", 0, "fp64"),
]


def _alone(spec):
    n, kind, jitter, seed, smoother, passes, precision = spec
    m = cube_mesh(n, kind, jitter, seed)
    gm = GpuMesh.from_cube(m)
    r = gm.smooth(smoother, passes, precision=precision, kernel="wave")
    pos = gm.positions()[gm.permutation()]       # back to the caller's numbering
    return pos, r


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _check(res, ref, tag):
    pos, r = ref
    assert _same(res["pos"], pos), tag
    assert _same(res["min"], r["min"]), tag
    assert _same(res["inv_sum"], r["inv_sum"]), tag
    assert np.array_equal(res["hist"], r["hist"]), tag


def test_stream_results_match_alone():
    refs = [_alone(s) for s in JOBS]
    with GpuServer(worker_threads=4) as srv:
        ids = []
        for k, (n, kind, jitter, seed, smoother, passes, precision) in enumerate(JOBS):
            ids.append(srv.submit_mesh(cube_mesh(n, kind, jitter, seed), smoother, passes, precision))
            time.sleep(0.002 * (k % 3))          # staggered arrivals: later jobs join running rounds
        for k, j in enumerate(ids):
            assert srv.wait(j, 60000) == "done", k
            _check(srv.result(j), refs[k], k)
        st = srv.stats()
        assert st["done"] == len(JOBS) and st["failed"] == 0
        assert st["peak_active"] >= 2            # jobs overlapped on the device


def test_texture_server_matches_alone():
    """A server created with the texture sampler (metric conformity, the
    shell field on 12^3 cells): jobs of every smoother, tets and hexes, come
    out as the same mesh smoothed alone under that crew; jobs before the
    texture is set are refused."""
    from gpumesh_amd.samplers import shell_texture
    specs = [(6, "hex", 0.3, 201, "nm", 2, "fp64"), (7, "tet", 0.3, 202, "gd", 3, "fp32_polish"),
             (5, "hex", 0.3, 203, "getme", 2, "fp64"), (6, "tet", 0.25, 204, "spawn", 2, "fp32"),
             (8, "hex", 0.3, 205, "gd", 2, "fp32_polish")]
    base = cube_mesh(4, "tet", 0.0, 1)
    tex = shell_texture(base.pos, 12, 3.0)
    refs = []
    for n, kind, jitter, seed, smoother, passes, precision in specs:
        m = cube_mesh(n, kind, jitter, seed)
        gm = GpuMesh.from_cube(m)
        gm.set_crew("metric_conformity", "texture", 3.0, tex)
        r = gm.smooth(smoother, passes, precision=precision, kernel="wave", spawn_coeff=1.0 / 20.0)
        refs.append((gm.positions()[gm.permutation()], r))
    with GpuServer(worker_threads=2, evaluator="metric_conformity", sampler="texture", scaling=3.0,
                   texture=tex) as srv:
        ids = [srv.submit_mesh(cube_mesh(n, kind, jitter, seed), smoother, passes, precision)
               for n, kind, jitter, seed, smoother, passes, precision in specs]
        for k, j in enumerate(ids):
            assert srv.wait(j, 60000) == "done", k
            _check(srv.result(j), refs[k], k)
    with GpuServer(worker_threads=1, evaluator="metric_conformity", sampler="texture", scaling=3.0) as srv:
        with pytest.raises(GpuMeshError):
            srv.submit_mesh(cube_mesh(4, "tet", 0.3, 1), "nm", 1, "fp64")


def test_admission_by_memory_budget():
    """A budget that holds about two of the jobs: jobs wait for memory, never
    exceed the budget, and still match their solo results; a job larger than
    the whole budget fails with GM_ERR_NOMEM instead of blocking the queue."""
    specs = [(12, "tet", 0.3, 200 + k, "getme", 3, "fp32_polish") for k in range(6)]
    refs = [_alone(s) for s in specs]
    probe = GpuMesh.from_cube(cube_mesh(12, "tet", 0.3, 200))
    probe.smooth("getme", 3, precision="fp32_polish", kernel="wave")
    one = probe.device_bytes()
    budget = int(2.6 * one)
    with GpuServer(budget_bytes=budget, worker_threads=2) as srv:
        big = cube_mesh(30, "tet", 0.3, 1)
        jb = srv.submit_mesh(big, "getme", 2, "fp32")
        ids = [srv.submit_mesh(cube_mesh(*s[:4]), s[4], s[5], s[6]) for s in specs]
        assert srv.wait(jb, 60000) == "failed"
        with pytest.raises(GpuMeshError) as e:
            srv.result(jb)
        assert e.value.status == 4
        for k, j in enumerate(ids):
            assert srv.wait(j, 60000) == "done"
            _check(srv.result(j), refs[k], k)
        st = srv.stats()
        assert st["peak_reserved_bytes"] <= budget
        assert st["peak_active"] <= 2


def test_slot_limit_and_release():
    with GpuServer(max_active=1, worker_threads=1) as srv:
        specs = [(5, "hex", 0.3, 300 + k, "nm", 2, "fp64") for k in range(3)]
        refs = [_alone(s) for s in specs]
        ids = [srv.submit_mesh(cube_mesh(*s[:4]), s[4], s[5], s[6]) for s in specs]
        for k, j in enumerate(ids):
            assert srv.wait(j) == "done"
            _check(srv.result(j), refs[k], k)
        assert srv.stats()["peak_active"] == 1
        with pytest.raises(GpuMeshError):
            srv.wait(ids[0])                     # released with its result


def test_job_files_match_alone():
    """Jobs submitted as job files (gm_server_submit_job) give the same bits
    as the same meshes submitted as arrays, and as smoothing them alone."""
    from gpumesh_amd import encode_job
    specs = JOBS[:6]
    refs = [_alone(s) for s in specs]
    with GpuServer(worker_threads=2) as srv:
        ids = [srv.submit_job(encode_job(cube_mesh(*s[:4]), s[4], s[5], s[6])) for s in specs]
        for k, j in enumerate(ids):
            assert srv.wait(j, 60000) == "done", k
            _check(srv.result(j), refs[k], k)
        bad = bytearray(encode_job(cube_mesh(3, "tet", 0.3, 1), "nm", 2))
        bad[-1] ^= 1
        with pytest.raises(GpuMeshError):
            srv.submit_job(bytes(bad))


def test_jobs_without_movable_vertices():
    """One-cell cubes have no movable vertex (no colour group).  Served
    first on their own (a round with no colour step), then next to ordinary
    jobs, each result equals the same mesh smoothed alone."""
    specs = [(1, "tet", 0.0, 1, "getme", 3, "fp32_polish"), (1, "hex", 0.0, 1, "gd", 2, "fp64"),
             (1, "tet", 0.0, 2, "nm", 2, "fp32"), (7, "tet", 0.3, 301, "getme", 3, "fp32_polish"),
             (6, "hex", 0.3, 302, "gd", 2, "fp64")]
    refs = [_alone(s) for s in specs]
    with GpuServer(worker_threads=2) as srv:
        first = [srv.submit_mesh(cube_mesh(*s[:4]), s[4], s[5], s[6]) for s in specs[:3]]
        for k, j in enumerate(first):
            assert srv.wait(j, 60000) == "done", k
            _check(srv.result(j), refs[k], k)
        ids = [srv.submit_mesh(cube_mesh(*s[:4]), s[4], s[5], s[6]) for s in specs]
        for k, j in enumerate(ids):
            assert srv.wait(j, 60000) == "done", k
            _check(srv.result(j), refs[k], k)
        assert srv.stats()["failed"] == 0


def test_shared_device_topology_and_prepare():
    """Meshes of one cached topology share its device arrays: results are
    bit-identical to uncached meshes, a mesh keeps working after another of
    its topology is destroyed, and a topology prepared ahead of time
    (gm_topology_prepare) serves later meshes the same way."""
    from gpumesh_amd import prepare_topology, set_topology_cache, topology_cache_stats
    ma, mb = cube_mesh(10, "tet", 0.3, seed=1), cube_mesh(10, "tet", 0.3, seed=2)
    mh = cube_mesh(9, "hex", 0.3, seed=3)
    ref = {}
    for key, m, sm in (("a", ma, "nm"), ("b", mb, "getme"), ("h", mh, "gd")):
        g = GpuMesh.from_cube(m)
        r = g.smooth(sm, 2, precision="fp64", kernel="wave")
        ref[key] = (g.positions(), r["min"])
        g.close()
    set_topology_cache(1 << 30)
    try:
        prepare_topology(mh.pos, mh.fixed, mh.tets, mh.hexs, device=True)
        a, b = GpuMesh.from_cube(ma), GpuMesh.from_cube(mb)
        st = topology_cache_stats()
        assert st["entries"] == 2 and st["hits"] >= 1
        ra = a.smooth("nm", 2, precision="fp64", kernel="wave")
        a.close()
        rb = b.smooth("getme", 2, precision="fp64", kernel="wave")
        assert _same(ra["min"], ref["a"][1]) and _same(rb["min"], ref["b"][1])
        assert _same(b.positions(), ref["b"][0])
        h = GpuMesh.from_cube(mh)
        rh = h.smooth("gd", 2, precision="fp64", kernel="wave")
        assert _same(h.positions(), ref["h"][0]) and _same(rh["min"], ref["h"][1])
        assert topology_cache_stats()["hits"] >= 2
        b.close()
        h.close()
    finally:
        set_topology_cache(0)
