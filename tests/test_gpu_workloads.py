"""The batched and served workloads at their own sizes.

Batched: 256 uneven cube meshes (n log-uniform in [8, 60], 30 % hexahedral),
GETMe, FP32, 10 passes replayed from a HIP graph after the autotuner chose
the launch shape; sampled members must be bit-identical to the same mesh
smoothed alone, and an FP64 run of the whole batch must give the oracle's
positions on its smallest members.

Served: 100 jobs from the serving stream's distribution (n log-uniform in
[8, 100], tet/hex, jitter 0.2-0.4 h, 5-20 passes, NM/GD/GETMe, FP32 with an
FP64 polish pass) arriving as a Poisson stream at 1000 jobs/s; every result
must be bit-identical to the same job run alone, and the reserved device
memory must stay within the budget.
"""
import time

import numpy as np
import pytest

from gpumesh_amd import GpuBatch, GpuMesh, GpuServer, cube_mesh, set_topology_cache

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _batched_specs(count=256):
    specs = []
    for s in range(1000, 1000 + count):
        rng = np.random.default_rng(s)
        n = int(round(np.exp(rng.uniform(np.log(8.0), np.log(60.0)))))
        specs.append((n, "hex" if rng.random() < 0.3 else "tet", s))
    return specs


@pytest.fixture
def topology_cache():
    set_topology_cache(48 << 30)
    try:
        yield
    finally:
        set_topology_cache(0)


def test_batched_256_meshes_match_alone_and_oracle(oracle, topology_cache):
    specs = _batched_specs()
    meshes = [cube_mesh(n, kind, 0.3, seed=s) for n, kind, s in specs]
    print(f"batched: {len(meshes)} meshes, {sum(m.nverts for m in meshes)} vertices", flush=True)
    gms = [GpuMesh.from_cube(m) for m in meshes]
    starts = [g.positions() for g in gms]
    b = GpuBatch(gms)
    (ppg, tile), _ = b.autotune("getme", precision="fp32", passes=1)
    assert all(np.array_equal(g.positions(), p) for g, p in zip(gms, starts))   # the autotuner restores
    b.smooth("getme", 10, precision="fp32", graph=True)                         # captures the pass graph
    for g, p in zip(gms, starts):
        g.set_positions(p)
    r = b.smooth("getme", 10, precision="fp32", graph=True)                     # replays it
    rng = np.random.default_rng(11)
    sample = set(rng.choice(len(gms), 13, replace=False).tolist())
    sample |= {int(np.argmax([m.nverts for m in meshes])), int(np.argmin([m.nverts for m in meshes])),
               next(k for k, s in enumerate(specs) if s[1] == "hex")}
    for k in sorted(sample):
        solo = GpuMesh.from_cube(meshes[k])
        rs = solo.smooth("getme", 10, precision="fp32", kernel="wave", counts=False)
        assert _same(gms[k].positions(), solo.positions()), (k, specs[k])
        assert _same(r["min"][k], rs["min"]) and _same(r["inv_sum"][k], rs["inv_sum"]), (k, specs[k])
        assert np.array_equal(r["hist"][k], rs["hist"]), (k, specs[k])
        solo.close()
    assert np.all(r["mean"][:, -1] > r["mean"][:, 0])
    # FP64: the whole batch, eagerly; the four smallest members against the oracle.
    for g, p in zip(gms, starts):
        g.set_positions(p)
    r64 = b.smooth("getme", 10, precision="fp64", graph=False)
    for k in np.argsort([m.nverts for m in meshes])[:4]:
        om = oracle.OracleMesh.from_cube(meshes[k])
        ref = om.smooth("getme", 10, threads=4)
        assert np.array_equal(gms[k].positions(), om.positions()), (int(k), specs[k])
        assert np.array_equal(r64["min"][k], ref["min"]), (int(k), specs[k])
    b.close()


def _served_jobs(count=100, seed=5):
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(8.0), np.log(100.0)))))
        kind = "tet" if rng.random() < 0.5 else "hex"
        jitter = float(rng.uniform(0.2, 0.4))
        passes = int(rng.integers(5, 21))
        smoother = ["nm", "gd", "getme"][int(rng.integers(0, 3))]
        jobs.append((n, kind, jitter, 5000 + k, smoother, passes))
    return jobs


def test_served_stream_at_1000_jobs_per_s_matches_alone(topology_cache):
    jobs = _served_jobs()
    meshes = [cube_mesh(n, kind, jit, seed=sd) for n, kind, jit, sd, _, _ in jobs]
    print(f"served: {len(jobs)} jobs, {sum(m.nverts for m in meshes)} vertices, n max "
          f"{max(j[0] for j in jobs)}", flush=True)
    budget = 288 * 10**9
    arrivals = np.cumsum(np.random.default_rng(1005).exponential(1.0 / 1000.0, len(jobs)))
    with GpuServer(budget_bytes=budget, worker_threads=8) as srv:
        ids = []
        t0 = time.perf_counter()
        for (n, kind, jit, sd, sm, p), m, at in zip(jobs, meshes, arrivals):
            delay = t0 + at - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit_mesh(m, sm, p, "fp32_polish"))
        results = []
        for k, j in enumerate(ids):
            assert srv.wait(j, 120000) == "done", (k, jobs[k])
            results.append(srv.result(j))
        st = srv.stats()
    print(f"served: span {time.perf_counter() - t0:.3f} s, peak active {st['peak_active']}, peak reserved "
          f"{st['peak_reserved_bytes'] / 1e9:.2f} GB", flush=True)
    assert st["done"] == len(jobs) and st["failed"] == 0
    assert 0 < st["peak_reserved_bytes"] <= budget
    assert st["peak_active"] >= 2
    for k, ((n, kind, jit, sd, sm, p), m) in enumerate(zip(jobs, meshes)):
        g = GpuMesh.from_cube(m)
        r = g.smooth(sm, p, precision="fp32_polish", kernel="wave", counts=False)
        pos = g.positions()[g.permutation()]
        res = results[k]
        assert _same(res["pos"], pos), (k, jobs[k])
        assert _same(res["min"], r["min"]) and _same(res["inv_sum"], r["inv_sum"]), (k, jobs[k])
        assert np.array_equal(res["hist"], r["hist"]), (k, jobs[k])
        g.close()
