"""Benchmark: smoothed vertices/s on the 1M-vertex Kuhn tetrahedral cube.

Workload (one GPU): the unit cube split into 100^3 cells, each Kuhn-split
into 6 tetrahedra (1,030,301 vertices, 6,000,000 tets, 970,299 movable),
interior vertices jittered by +-0.35 h (seed 2, no inverted element), mean
ratio evaluator with the metric-free measurer.  Gradient descent, GETMe and
spawn search (64 lattice offsets scaled by 0.25 x the local edge length) each
smooth the mesh from the same start for K passes in the mixed-precision mode
(FP32 passes, the last pass an FP64 polish, every per-pass min/mean reduction
in FP64).  A step is one smoothing pass of one smoother over the whole mesh.

value = movable vertices x passes x smoothers / wall time of the timed passes
(inputs resident in HBM; the per-pass FP64 quality reductions are inside the
timed region).  The timed passes replay each pass body from a HIP graph (the
mesh as a batch of one); an untimed eager run of the same passes, which must
give identical results, supplies the per-launch kernel times.  Also reported:
  * the plain FP64 single-mesh kernel (one lane per vertex) on the same
    workload, the denominator of the speed-up;
  * the exact FP64 kernels (the reference's operation order) and the
    mixed-precision FP64 kernels for every pass (fp64_fast), with each run's
    final min/mean against the exact FP64 run of the same passes; fp64_fast
    is the mode whose quality agrees within 1e-6 ("equal_quality");
  * the roofline of the dominant kernel: the work its formulation performs
    per launch over its average launch time (HIP events), with the HBM bytes
    of its committed rocprofv3 profile beside it;
  * the CPU oracle (the reference's Thread mode on the workload mesh, Serial
    mode on one pinned core on a smaller mesh).

Two more legs run after it and are reported beside the headline value:
  * batched: 256 meshes (Kuhn tet or, with probability 0.3, hex cubes of n^3
    cells, n log-uniform in [8, 60], seeds 1000..1255, jitter 0.3 h) smoothed
    together by GETMe for 10 FP32 passes, one fused launch per colour step and
    element kind, each pass replayed from a HIP graph, launch shape chosen by
    the autotuner; compared with the same meshes smoothed one by one;
  * serving: a stream of jobs (tet or hex cubes, n log-uniform in [8, 100],
    jitter U[0.2, 0.4] h, 5-20 passes, Nelder-Mead, gradient descent or
    GETMe, FP32 with an FP64 polish pass) arriving as a Poisson process at
    several rates, admitted by free device memory and batched continuously;
    p50/p99 latency, jobs/s and vertices/s per rate.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 2]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Peak rates of one MI355X (AMD published figures).
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0

# Algorithmic FLOP, counting +, -, *, /, sqrt and pow as one each.  One
# mean-ratio tetrahedron evaluated inside a patch: 3 metric-free edge vectors
# 18, F = E Fr^-1 45, det 14, |F|^2 17, corner quality 4 (incl. pow), patch
# accumulation 3 (1/q and two adds).
FLOP_MR_TET_PATCH = 101.0
# One GETMe tetrahedron transform: 4 face normals 60, 4 scaled moves 56,
# 2 volumes 50, volume ratio and cube root 3, centre 12, rescale 36,
# 2 qualities 196, weight 2.
FLOP_GETME_TET = 415.0
# GETMe vertex update, per patch element pulled: 3 mul + 4 add.
FLOP_GETME_PULL = 7.0
# FLOP the FP32 kernels execute per tetrahedron evaluation on the affine
# corner terms (the formulation's own work, the unit of the roofline below):
# determinant d0 + g.x 6 and squared norm c0 + h.x + s|x|^2 8 (these 14 on the
# matrix cores in the spawn-search and gradient-descent kernels), |d|^(-2/3)
# and 3/q 4, accumulation 2.
FLOP_AFFINE_TET = 20.0
# HBM bytes per launch of a kernel come from its committed rocprofv3 --pmc
# summary (FETCH_SIZE and WRITE_SIZE, scaled by the same counters' reading for
# a 1 GiB device copy); the summary is regenerated whenever the kernel changes.
TRAFFIC_PROFILES = {("spawnMfmaKernel<0, 4>", 100): "profiles/spawn_mfma_tet100_hbm.md"}


def issue_rates():
    """FP32 vector issue rates of the device (gm_throughput_probe): chains of
    the log2/exp2 pair, chains of FMAs, and both in one loop."""
    from gpumesh_amd._lib import check, lib
    r = [np.zeros(3) for _ in range(3)]
    for op in range(3):
        check(lib().gm_throughput_probe(op, r[op].ctypes.data), "gm_throughput_probe")
    return {"trans_lane_ops_per_s": float(r[0][0]), "fma_lane_ops_per_s": float(r[1][1]),
            "trans_ms": float(r[0][2]), "fma_ms": float(r[1][2]), "mixed_ms": float(r[2][2])}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(kernel, n):
    """(bytes per launch, source) from the kernel's committed HBM profile, or (None, reason)."""
    import re
    rel = TRAFFIC_PROFILES.get((kernel, n))
    if rel is None:
        return None, "no HBM profile for this kernel and size"
    try:
        text = open(os.path.join(ROOT, rel)).read()
    except OSError:
        return None, f"{rel} not present"
    m = re.search(r"HBM traffic per launch: [0-9.]+ MB \((\d+) bytes\)", text)
    if not m:
        return None, f"{rel} holds no traffic line"
    return float(m.group(1)), f"{rel} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; profiled separately, not this run)"


def patch_sizes(mesh):
    """Elements in each vertex's patch (new numbering, the order of eval_counts)."""
    eoff = mesh.topology()[0].astype(np.int64)
    return np.diff(eoff)


def evaluated_elements(res, sizes):
    """Element evaluations of a run: each vertex's patch evaluations times its patch size."""
    return float(np.dot(res["eval_counts"].astype(np.float64), sizes))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, passes, threads, smoothers, spawn_coeff, cores=None):
    """The FP64 CPU oracle (the reference's Thread mode with `threads` workers,
    glibc libm) on a bounded sample: `passes` sweeps of each smoother on the
    n^3 workload construction.  `cores` pins the process to those cores for
    the timed runs (the reference's Serial mode runs on one)."""
    from gpumesh_amd import cube_mesh
    from oracle import pyoracle
    pyoracle.set_libm(0)
    m = cube_mesh(n, "tet", 0.35, seed=2)
    moved = 0
    t_total = 0.0
    saved = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else None
    try:
        if cores is not None and saved is not None:
            os.sched_setaffinity(0, cores)
        for sm in smoothers:
            om = pyoracle.OracleMesh.from_cube(m)
            t0 = time.perf_counter()
            om.smooth(sm, passes, threads=threads, order=1 if threads > 1 else 0, eval_chunks=threads,
                      spawn_coeff=spawn_coeff)
            t_total += time.perf_counter() - t0
            moved += m.nmovable * passes
    finally:
        if cores is not None and saved is not None:
            os.sched_setaffinity(0, saved)
        pyoracle.set_libm(1)
    return moved / t_total, t_total, m


def batched_leg(args):
    """256 uneven meshes, GETMe, FP32, fused launches, graph replay, autotuned."""
    from gpumesh_amd import GpuBatch, GpuMesh, cube_mesh
    specs = []
    for s in range(1000, 1000 + args.batch_meshes):
        rng = np.random.default_rng(s)
        n = int(round(np.exp(rng.uniform(np.log(8.0), np.log(60.0)))))
        specs.append((n, "hex" if rng.random() < 0.3 else "tet", s))
    t0 = time.perf_counter()
    meshes = [cube_mesh(n, kind, 0.3, seed=s) for n, kind, s in specs]
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    gms = [GpuMesh.from_cube(m) for m in meshes]
    t_up = time.perf_counter() - t0
    starts = [g.positions() for g in gms]
    movable = sum(g.nmovable for g in gms)
    elems = sum(g.ntets + g.nhexs for g in gms)
    passes = args.batch_passes
    log(f"batched: {len(gms)} meshes, {sum(m.nverts for m in meshes)} vertices ({movable} movable), {elems} elements, "
        f"{sum(1 for _, k, _ in specs if k == 'hex')} hex meshes; generate {t_gen:.2f}s, topology+upload {t_up:.2f}s")

    def reset():
        for g, p in zip(gms, starts):
            g.set_positions(p)

    b = GpuBatch(gms)
    (ppg, tile), times = b.autotune("getme", precision="fp32", passes=1)
    log(f"batched autotune getme: best {ppg} wavefronts/block; sweep ms (tile 64 column) "
        f"{[round(float(x), 3) for x in times[:, 0]]}")
    b.smooth("getme", passes, precision="fp32", graph=True)          # captures the pass graph
    reset()
    t0 = time.perf_counter()
    r = b.smooth("getme", passes, precision="fp32", graph=True)
    wall = time.perf_counter() - t0
    replayed = [g.positions() for g in gms]
    reset()
    t0 = time.perf_counter()
    re = b.smooth("getme", passes, precision="fp32", graph=False)
    wall_eager = time.perf_counter() - t0
    # Graph replay, the eager run and every member alone must agree bit for bit.
    same = lambda a, b: np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
    if not (all(same(g.positions(), p) for g, p in zip(gms, replayed)) and same(r["inv_sum"], re["inv_sum"])
            and same(r["min"], re["min"])):
        raise RuntimeError("batched leg: the graph-replayed passes differ from the eager ones")
    # The same meshes one by one (FP32, then the plain FP64 kernel).
    reset()
    t0 = time.perf_counter()
    solo = [g.smooth("getme", passes, precision="fp32", kernel="wave", counts=False) for g in gms]
    wall_single = time.perf_counter() - t0
    if not all(same(g.positions(), p) and same(sr["inv_sum"], r["inv_sum"][i])
               for i, (g, p, sr) in enumerate(zip(gms, replayed, solo))):
        raise RuntimeError("batched leg: a member differs from the same mesh smoothed alone")
    reset()
    t0 = time.perf_counter()
    for g in gms:
        g.smooth("getme", passes, precision="fp64", kernel="plain", counts=False)
    wall_plain = time.perf_counter() - t0
    log(f"batched getme x{passes}: graph {wall * 1e3:.1f} ms, eager {wall_eager * 1e3:.1f} ms, one by one "
        f"{wall_single * 1e3:.1f} ms, plain fp64 one by one {wall_plain * 1e3:.1f} ms")
    b.close()
    return {"workload": f"{len(gms)} cube meshes, n log-uniform [8,60], 30% hex, seeds 1000..{999 + len(gms)}, "
                        f"jitter 0.3h, mean ratio, GETMe x {passes} passes, fp32, graph-replayed fused passes",
            "meshes": len(gms), "movable_vertices": movable, "elements": elems, "passes": passes,
            "vertices_per_s": movable * passes / wall, "ms_per_pass": wall / passes * 1e3,
            "elements_per_s": elems * passes / wall,
            "eager_vertices_per_s": movable * passes / wall_eager,
            "one_by_one_fp32_vertices_per_s": movable * passes / wall_single,
            "one_by_one_plain_fp64_vertices_per_s": movable * passes / wall_plain,
            "speedup_vs_plain_fp64": wall_plain / wall,
            "autotune": {"best_patches_per_group": ppg, "sweep_ms_by_ppg": [float(x) for x in times[:, 0]]},
            "checked": "graph replay == eager run == every member alone (positions and pass records, bit for bit)",
            "final_min_mean": [float(np.min(r["min"][:, -1])), float(np.mean(r["mean"][:, -1]))]}


def topology_leg(args):
    """The reference's global loop with topology operations on: per global
    pass, the BATR restructure on the host (gm_mesh_restructure) and
    relocation passes on the device."""
    from gpumesh_amd import GpuMesh, cube_mesh
    n = args.topo_n
    m = cube_mesh(n, "tet", 0.35, seed=2)
    gm = GpuMesh.from_cube(m)
    gm.set_crew("mean_ratio", "uniform", scaling=float(n))
    q0 = gm.evaluate()
    passes = []
    for g in range(args.topo_global):
        tets_in = gm.ntets
        t0 = time.perf_counter()
        counts = gm.restructure(5, 5, 0.5, 1.5)
        t_restructure = time.perf_counter() - t0
        t0 = time.perf_counter()
        r = gm.smooth("gd", args.topo_reloc, precision="fp32_polish", kernel="wave", counts=False)
        t_reloc = time.perf_counter() - t0
        passes.append({"tets_in": tets_in, "counts": counts, "restructure_s": t_restructure,
                       "tets_per_s": tets_in / t_restructure, "relocation_wall_ms": t_reloc * 1e3,
                       "relocation_device_ms": r["total_ms"], "vertices": gm.nverts, "tets": gm.ntets,
                       "min": float(r["min"][-1]), "mean": float(r["mean"][-1])})
        log(f"topology pass {g}: restructure {t_restructure:.2f} s ({counts}), {gm.nverts} vertices "
            f"{gm.ntets} tets; {args.topo_reloc} GD passes {t_reloc * 1e3:.1f} ms wall; "
            f"min {passes[-1]['min']:.4f} mean {passes[-1]['mean']:.4f}")
    return {"workload": f"Kuhn tet cube {n}^3 cells, jitter 0.35h seed 2, mean ratio (K = {n}), "
                        f"{args.topo_global} global passes of BATR restructure (5 topology passes, 5 refinement "
                        f"sweeps, edges in [0.5, 1.5]) + {args.topo_reloc} GD passes fp32_polish",
            "initial_min_mean": [q0["min"], q0["mean"]], "passes": passes}


def aniso_leg(args):
    """The anisotropic hexahedral workload (64^3 hexes, jitter 0.3 h, seed 3,
    metric conformity with the metric-wise measurer, K = 64, Nelder-Mead x 15)
    under the analytic shell field and its 64^3 FP32 texture: FP32 with the
    FP64 polish, exact FP64 (wave kernel), and the plain FP64 kernel (2
    passes), wall clock around each run after a one-pass warm-up."""
    from gpumesh_amd import GpuMesh, cube_mesh
    from gpumesh_amd.samplers import shell_texture
    m = cube_mesh(64, "hex", 0.3, seed=3)
    out = {"workload": "hex cube 64^3 cells (274625 vertices), jitter 0.3h seed 3, metric conformity, K = 64, "
                       "Nelder-Mead x 15", "runs": {}}
    for sampler in ("analytic", "texture"):
        tex = shell_texture(m.pos, 64, 64.0) if sampler == "texture" else None
        for prec, kernel, passes in (("fp32_polish", "wave", 15), ("fp64", "wave", 15), ("fp64", "plain", 2)):
            def fresh():
                g = GpuMesh.from_cube(m)
                g.set_crew("metric_conformity", sampler, scaling=64.0, texture=tex)
                return g
            fresh().smooth("nm", 1, precision=prec, kernel=kernel)
            gm = fresh()
            t0 = time.perf_counter()
            r = gm.smooth("nm", passes, precision=prec, kernel=kernel)
            wall = time.perf_counter() - t0
            key = f"{sampler}/{prec}/{kernel}"
            out["runs"][key] = {"passes": passes, "wall_ms": wall * 1e3, "vertices_per_s": gm.nmovable * passes / wall,
                                "final_min": float(r["min"][-1]), "final_mean": float(r["mean"][-1])}
            log(f"aniso {key}: {passes} passes {wall * 1e3:.1f} ms, {gm.nmovable * passes / wall:.3g} vertices/s")
    return out


def serving_leg(args):
    """Poisson job stream through the serving loop at several rates.  Every
    (kind, n) shape of the job list is built into the topology cache, with
    its device arrays, before the stream (untimed, reported), as a server
    that knows its shape catalogue does; jobs of one shape share the element
    arrays."""
    from concurrent.futures import ThreadPoolExecutor

    from gpumesh_amd import GpuMesh, GpuServer, cube_mesh, prepare_topology, set_topology_cache, topology_cache_stats
    rng = np.random.default_rng(5)
    jobs = []
    for k in range(args.serve_jobs):
        n = int(round(np.exp(rng.uniform(np.log(8.0), np.log(100.0)))))
        kind = "tet" if rng.random() < 0.5 else "hex"
        jitter = float(rng.uniform(0.2, 0.4))
        passes = int(rng.integers(5, 21))
        smoother = ["nm", "gd", "getme"][int(rng.integers(0, 3))]
        jobs.append((n, kind, jitter, 5000 + k, smoother, passes))
    shapes = sorted({(kind, n) for n, kind, _, _, _, _ in jobs})
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    base = {}
    with ThreadPoolExecutor(threads) as ex:
        for (kind, n), m in zip(shapes, ex.map(lambda s: cube_mesh(s[1], s[0], 0.0, seed=1), shapes)):
            base[(kind, n)] = m

        def make(j):
            n, kind, jit, sd, _, _ = j
            m = cube_mesh(n, kind, jit, seed=sd)
            b = base[(kind, n)]
            m.tets, m.hexs = b.tets, b.hexs            # one element array per shape
            return m
        meshes = list(ex.map(make, jobs))
    t_gen = time.perf_counter() - t0
    log(f"serving: {len(jobs)} jobs, {len(shapes)} shapes, {sum(m.nverts for m in meshes)} vertices in total, "
        f"generated in {t_gen:.2f}s")
    set_topology_cache(args.serve_cache_gb << 30)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda k: prepare_topology(base[k].pos, base[k].fixed, base[k].tets, base[k].hexs, device=True),
                    shapes))
    t_prep = time.perf_counter() - t0
    cache = topology_cache_stats()
    log(f"serving: {cache['entries']} topologies prepared (host and device) in {t_prep:.2f}s on {threads} threads, "
        f"{cache['bytes'] / 1e9:.1f} GB host")
    work = sum(m.nmovable * j[5] for m, j in zip(meshes, jobs))
    rates = [float(x) for x in args.serve_rates.split(",") if x]
    out = {"workload": f"{len(jobs)} jobs per rate: tet/hex 50/50 cubes, n log-uniform [8,100], jitter U[0.2,0.4]h, "
                       f"5-20 passes, NM/GD/GETMe uniform, fp32 + fp64 polish, Poisson arrivals",
           "jobs": len(jobs), "vertex_passes_per_run": work, "shapes": len(shapes),
           "prepare_s": t_prep, "prepare_threads": threads, "cache_host_gb": cache["bytes"] / 1e9, "rates": {}}
    with GpuServer(budget_bytes=288 * 10**9, worker_threads=args.serve_workers) as srv:
        # Untimed warm-up jobs (first launches of each kernel).
        warm = [srv.submit_mesh(meshes[0], sm, 1) for sm in ("nm", "gd", "getme")]
        for w in warm:
            srv.wait(w)
            srv.release(w)
        for rate in rates:
            st0 = srv.stats()
            arr = np.cumsum(np.random.default_rng(int(rate) + 5).exponential(1.0 / rate, len(jobs)))
            ids = []
            t0 = time.perf_counter()
            for (n, kind, jit, sd, sm, p), m, at in zip(jobs, meshes, arr):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit_mesh(m, sm, p, "fp32_polish"))
            t_sub = time.perf_counter() - t0
            lat, parts = [], []
            for j in ids:
                st = srv.wait(j, 120000)
                if st != "done":
                    raise RuntimeError(f"serving job ended {st}")
                r = srv.result(j)
                lat.append(r["latency_ms"])
                parts.append((r["build_ms"], r["admit_wait_ms"], r["device_ms"]))
            span = time.perf_counter() - t0
            st1 = srv.stats()
            lat = np.array(lat)
            parts = np.array(parts)
            out["rates"][str(int(rate))] = {
                "offered_jobs_per_s": rate, "jobs_per_s": len(jobs) / span, "vertices_per_s": work / span,
                "arrival_span_s": float(arr[-1]), "submit_span_s": t_sub, "span_s": span,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "max_ms": float(lat.max()),
                "p50_build_ms": float(np.percentile(parts[:, 0], 50)),
                "p99_build_ms": float(np.percentile(parts[:, 0], 99)),
                "p50_admit_wait_ms": float(np.percentile(parts[:, 1], 50)),
                "p50_device_ms": float(np.percentile(parts[:, 2], 50)),
                "rounds": st1["rounds"] - st0["rounds"],
                "round_busy_frac": (st1["round_ms"] - st0["round_ms"]) / 1e3 / span}
            log(f"serving rate {rate:.0f}/s: {len(jobs) / span:.1f} jobs/s over {span:.2f}s (arrivals span "
                f"{arr[-1]:.2f}s), {work / span / 1e6:.1f} M vertices/s, p50 {np.percentile(lat, 50):.1f} ms, "
                f"p99 {np.percentile(lat, 99):.1f} ms; build p50 {np.percentile(parts[:, 0], 50):.2f} ms; rounds "
                f"{st1['rounds'] - st0['rounds']:.0f}, busy {(st1['round_ms'] - st0['round_ms']) / 1e3 / span:.2f}")
        st = srv.stats()
        out["server"] = {"rounds": st["rounds"], "round_wall_s": st["round_ms"] / 1e3, "peak_active": st["peak_active"],
                         "peak_reserved_gb": st["peak_reserved_bytes"] / 1e9, "budget_gb": st["budget_bytes"] / 1e9}
    # The same jobs one by one through the plain FP64 kernel (a bounded sample).
    sample = list(range(min(args.serve_plain_sample, len(jobs))))
    t_plain = 0.0
    w_plain = 0
    for k in sample:
        g = GpuMesh.from_cube(meshes[k])
        t0 = time.perf_counter()
        g.smooth(jobs[k][4], jobs[k][5], precision="fp64", kernel="plain", counts=False)
        t_plain += time.perf_counter() - t0
        w_plain += meshes[k].nmovable * jobs[k][5]
        g.close()
    out["plain_fp64_one_by_one_vertices_per_s"] = w_plain / t_plain if t_plain else None
    out["plain_sample_jobs"] = len(sample)
    log(f"serving plain fp64 sample: {len(sample)} jobs, {w_plain / max(t_plain, 1e-9) / 1e6:.1f} M vertices/s")
    set_topology_cache(0)            # drops the serving shapes and their device copies
    set_topology_cache(8 << 30)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed passes per smoother")
    ap.add_argument("--warmup", type=int, default=2, help="untimed passes per smoother before timing")
    ap.add_argument("--n", type=int, default=100, help="cells per cube edge")
    ap.add_argument("--smoothers", default="gd,getme,spawn")
    ap.add_argument("--precision", default="fp32_polish", choices=["fp64", "fp32", "fp32_polish", "fp64_fast"])
    ap.add_argument("--polish-passes", type=int, default=1, help="FP64 passes ending an fp32_polish run")
    ap.add_argument("--kernel", default="wave", choices=["plain", "wave"])
    ap.add_argument("--plain-steps", type=int, default=2, help="passes of the plain FP64 kernel per smoother")
    ap.add_argument("--cpu-n", type=int, default=100, help="cube size of the CPU Thread-mode sample")
    ap.add_argument("--cpu-passes", type=int, default=1, help="passes per smoother of the CPU sample")
    ap.add_argument("--cpu-serial-n", type=int, default=40, help="cube size of the CPU Serial-mode sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-meshes", type=int, default=256)
    ap.add_argument("--batch-passes", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=1000)
    ap.add_argument("--serve-cache-gb", type=int, default=96, help="topology cache of the serving leg (host GB)")
    ap.add_argument("--serve-rates", default="50,200,1000,2000")
    ap.add_argument("--serve-plain-sample", type=int, default=8)
    ap.add_argument("--serve-workers", type=int, default=8, help="host worker threads of the server")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--topo-n", type=int, default=20, help="cube size of the topology leg")
    ap.add_argument("--topo-global", type=int, default=2, help="global passes of the topology leg")
    ap.add_argument("--topo-reloc", type=int, default=5, help="relocation passes per global pass")
    ap.add_argument("--no-topo", action="store_true")
    ap.add_argument("--aniso-only", action="store_true",
                    help="run only the anisotropic hexahedral workload's timings and print their JSON")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving leg and print its JSON")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 5 times (the reference kernel for HBM counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "repeats": 5}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(meshes shard with no exchange; see the sharding design notes).", file=sys.stderr)
        sys.exit(2)

    import gpumesh_amd
    from gpumesh_amd import GpuBatch, GpuMesh, cube_mesh

    if gpumesh_amd.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    smoothers = [s for s in args.smoothers.split(",") if s]
    spawn_coeff = 0.25
    issue = issue_rates()
    log(f"issue rates: transcendental {issue['trans_lane_ops_per_s'] / 1e12:.2f} T lane-ops/s, FMA "
        f"{issue['fma_lane_ops_per_s'] / 1e12:.2f} T lane-ops/s, both {issue['mixed_ms']:.3f} ms vs "
        f"{issue['trans_ms'] + issue['fma_ms']:.3f} ms apart")
    from gpumesh_amd._lib import check, lib
    check(lib().gm_set_topology_cache(8 << 30), "gm_set_topology_cache")
    check(lib().gm_set_polish_passes(args.polish_passes), "gm_set_polish_passes")
    if args.serve_only:
        print(json.dumps(serving_leg(args)))
        return
    if args.aniso_only:
        print(json.dumps(aniso_leg(args)))
        return

    t0 = time.perf_counter()
    m = cube_mesh(args.n, "tet", 0.35, seed=2)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    gm = GpuMesh.from_cube(m)
    t_up = time.perf_counter() - t0
    log(f"mesh n={args.n}: {m.nverts} vertices, {m.tets.shape[0]} tets, {gm.nmovable} movable, "
        f"{gm.ngroups} colour groups; generate {t_gen:.2f}s, topology+upload {t_up:.2f}s")
    start = gm.positions()

    per = {}
    # The timed passes replay captured pass bodies (hipGraph) on the mesh as a
    # batch of one; the warm-up runs at least two passes so that the FP32 and
    # the FP64 (polish) bodies are captured before the timed region.  The
    # same passes run once more eagerly, untimed, with per-launch HIP events
    # (kernel times for the roofline) and per-vertex evaluation counts; their
    # results must be identical bit for bit.
    one = GpuBatch([gm])
    for sm in smoothers:
        gm.set_positions(start)
        one.smooth(sm, max(args.warmup, 2), precision=args.precision, kernel=args.kernel, spawn_coeff=spawn_coeff,
                   graph=True)
        gm.set_positions(start)
        t0 = time.perf_counter()
        rg = one.smooth(sm, args.steps, precision=args.precision, kernel=args.kernel, spawn_coeff=spawn_coeff,
                        graph=True)
        wall = time.perf_counter() - t0
        end_pos = gm.positions()
        gm.set_positions(start)
        res = gm.smooth(sm, args.steps, precision=args.precision, kernel=args.kernel, spawn_coeff=spawn_coeff)
        if not (np.array_equal(end_pos, gm.positions()) and np.array_equal(rg["inv_sum"][0], res["inv_sum"])
                and np.array_equal(rg["min"][0], res["min"])):
            raise RuntimeError(f"{sm}: graph replay and eager run differ")
        res["graph_total_ms"] = rg["total_ms"]
        per[sm] = {"wall_s": wall, "res": res}
        log(f"{sm}: {args.steps} passes {wall * 1e3:.1f} ms wall ({res['graph_total_ms']:.1f} ms device; eager "
            f"{res['total_ms']:.1f} ms), "
            f"min {res['min'][0]:.4f} -> {res['min'][-1]:.4f}, mean {res['mean'][0]:.5f} -> {res['mean'][-1]:.5f}")

    total_wall = sum(p["wall_s"] for p in per.values())
    moved = gm.nmovable * args.steps * len(smoothers)
    value = moved / total_wall

    # Plain FP64 single-mesh kernel on the same workload (the baseline the
    # mixed-precision batched path is compared with).
    plain = {}
    for sm in smoothers:
        if args.plain_steps <= 0:
            break
        gm.set_positions(start)
        t0 = time.perf_counter()
        r = gm.smooth(sm, args.plain_steps, precision="fp64", kernel="plain", spawn_coeff=spawn_coeff)
        plain[sm] = {"wall_s": time.perf_counter() - t0, "passes": args.plain_steps, "res": r}
        log(f"plain fp64 {sm}: {args.plain_steps} passes {plain[sm]['wall_s'] * 1e3:.1f} ms")
    plain_value = None
    if plain:
        plain_value = sum(gm.nmovable * p["passes"] for p in plain.values()) / sum(p["wall_s"] for p in plain.values())

    # Dominant kernel class of the timed runs, and its roofline.  `achieved`
    # counts the work the kernel's formulation performs: for the affine FP32
    # smoothing kernels FLOP_AFFINE_TET per tetrahedron evaluation (element
    # evaluations counted exactly: each vertex's patch evaluations times its
    # patch size), for the evaluation pass its compulsory bytes.  The rate of
    # the reference's direct formula (FLOP_MR_TET_PATCH per evaluation) is
    # reported beside it as a work-reduction ratio, not as the fraction.
    sizes = patch_sizes(gm)
    best = None
    for sm, p in per.items():
        res = p["res"]
        km, kn = res["kernel_ms"], res["kernel_launches"]
        for cls in ("smooth_f32", "getme_elem_f32", "getme_vert_f32", "evaluate"):
            if kn[cls] and (best is None or km[cls] > best["ms"]):
                best = {"smoother": sm, "class": cls, "ms": km[cls], "launches": kn[cls], "res": res}
    roofline = None
    if best is not None:
        avg_s = best["ms"] / best["launches"] / 1e3
        steps = args.steps
        f32_passes = steps - min(args.polish_passes, steps) if args.precision == "fp32_polish" else steps
        if best["class"] == "smooth_f32":
            # The FP32 share of the counted evaluations (spawn search makes the
            # same 128 per vertex in every pass).
            elems = evaluated_elements(best["res"], sizes) * (f32_passes / steps) / best["launches"]
            kname = {"spawn": "spawnMfmaKernel<0, 4>", "gd": "gdMfmaTetKernel<3>",
                     "nm": "nmQuadKernel<6>"}.get(best["smoother"], "smoothWaveKernel<float>")
            flop = elems * FLOP_AFFINE_TET
            achieved = flop / avg_s / 1e12
            traffic, source = profiled_traffic(kname, args.n)
            roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "traffic_source": source,
                        "kernel": kname, "avg_launch_ms": avg_s * 1e3, "flop_per_launch": flop,
                        "flop_unit": f"{FLOP_AFFINE_TET:.0f} FLOP per tetrahedron evaluation on affine corner terms "
                                     f"({elems:.4g} element evaluations per launch)",
                        "peak_note": "157.3 = FP32 matrix-core / packed-vector peak; the vector ALUs run "
                                     "non-packed FP32 here (78.6 TFLOP/s)",
                        "reference_formula_tflops": elems * FLOP_MR_TET_PATCH / avg_s / 1e12,
                        "work_reduction_vs_reference_formula": FLOP_MR_TET_PATCH / FLOP_AFFINE_TET}
            if traffic:
                roofline["hbm_gbs"] = traffic / avg_s / 1e9
                roofline["hbm_frac"] = roofline["hbm_gbs"] / PEAK_HBM_GBS
            # The vector-issue ceiling measured on this device: every element
            # evaluation issues a log2 and an exp2 (transcendental rate) and a
            # scale and an accumulate (FMA rate); the affine terms run on the
            # matrix cores.  The two kinds share issue slots (the probe's mixed
            # run takes the sum of the two times).
            trans, fma = issue["trans_lane_ops_per_s"], issue["fma_lane_ops_per_s"]
            core_s = elems * (2.0 / trans + 2.0 / fma)
            roofline["issue_ceiling"] = dict(issue, core_instructions_per_evaluation="2 transcendental + 2 FMA-rate",
                                             evaluations_per_launch=elems, core_ms_per_launch=core_s * 1e3,
                                             frac=core_s / avg_s)
        elif best["class"] == "evaluate":
            bytes_per = 16.0 * gm.ntets + 32.0 * gm.nverts
            achieved = bytes_per / avg_s / 1e9
            roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": "evalKernel",
                        "avg_launch_ms": avg_s * 1e3, "bytes_per_launch": bytes_per}
        else:
            if best["class"] == "getme_elem_f32":
                flop_per_launch = FLOP_GETME_TET * gm.ntets
                kname = "getmeTetKernel<float, 0, 0>"
            else:
                flop_per_launch = (2 * FLOP_MR_TET_PATCH + FLOP_GETME_PULL) * float(np.sum(sizes[gm.nverts - gm.nmovable:])) \
                    / best["launches"] * f32_passes
                kname = "getmeWaveKernel<float, 0, 0, 32, 0>"
            achieved = flop_per_launch / avg_s / 1e12
            roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": kname,
                        "avg_launch_ms": avg_s * 1e3, "flop_per_launch": flop_per_launch}
        if roofline["frac"] > 1.0:
            raise RuntimeError(f"roofline fraction {roofline['frac']:.3f} above 1: the work unit is wrong")

    # FP64 on the same workload: the exact kernels (the reference's operation
    # order, bit-identical to the plain FP64 kernel and the CPU oracle) and the
    # mixed-precision FP64 kernels for every pass (GM_FP64_FAST).  The final
    # quality of each run is compared with the exact FP64 run of the same
    # passes; "equal quality" means final min and mean within 1e-6 of it.
    quality, fp64_exact, fp64_fast = {}, {}, {}
    for sm, p in per.items():
        gm.set_positions(start)
        t0 = time.perf_counter()
        r64 = gm.smooth(sm, args.steps, precision="fp64", kernel="wave", spawn_coeff=spawn_coeff, counts=False)
        fp64_exact[sm] = {"wall_s": time.perf_counter() - t0, "device_ms": r64["total_ms"]}
        gm.set_positions(start)
        one.smooth(sm, max(args.warmup, 1), precision="fp64_fast", spawn_coeff=spawn_coeff, graph=True)
        gm.set_positions(start)
        t0 = time.perf_counter()
        rf = one.smooth(sm, args.steps, precision="fp64_fast", spawn_coeff=spawn_coeff, graph=True)
        fp64_fast[sm] = {"wall_s": time.perf_counter() - t0, "device_ms": rf["total_ms"],
                         "min_diff": float(rf["min"][0][-1] - r64["min"][-1]),
                         "mean_diff": float(rf["mean"][0][-1] - r64["mean"][-1])}
        quality[sm] = {"fp64_final_min": float(r64["min"][-1]), "fp64_final_mean": float(r64["mean"][-1]),
                       "min_diff": float(p["res"]["min"][-1] - r64["min"][-1]),
                       "mean_diff": float(p["res"]["mean"][-1] - r64["mean"][-1])}
        log(f"{sm}: final min/mean vs fp64: {args.precision} {quality[sm]['min_diff']:+.2e} / "
            f"{quality[sm]['mean_diff']:+.2e}; fp64_fast {fp64_fast[sm]['min_diff']:+.2e} / "
            f"{fp64_fast[sm]['mean_diff']:+.2e}; exact fp64 {fp64_exact[sm]['wall_s'] * 1e3:.1f} ms, "
            f"fp64_fast {fp64_fast[sm]['wall_s'] * 1e3:.1f} ms wall")
    fast_value = moved / sum(v["wall_s"] for v in fp64_fast.values())
    exact_value = moved / sum(v["wall_s"] for v in fp64_exact.values())
    worst_min = max(abs(v["min_diff"]) for v in fp64_fast.values())
    worst_mean = max(abs(v["mean_diff"]) for v in fp64_fast.values())
    equal_quality = {
        "precision": "fp64_fast", "vertices_per_s": fast_value,
        "speedup_vs_plain_fp64": (fast_value / plain_value) if plain_value else None,
        "max_abs_final_min_diff": worst_min, "max_abs_final_mean_diff": worst_mean,
        "within_1e-6": bool(worst_min <= 1e-6 and worst_mean <= 1e-6),
        "per_smoother": fp64_fast}
    headline_equal = max(max(abs(q["min_diff"]), abs(q["mean_diff"])) for q in quality.values()) <= 1e-6

    batched = None if args.no_batch else batched_leg(args)
    serving = None if args.no_serve else serving_leg(args)
    topology = None if args.no_topo else topology_leg(args)

    cpu = None
    if not args.no_cpu:
        avail = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
        threads = max(1, min(len(avail), 16))
        cpu_value, cpu_s, cm = cpu_baseline(args.cpu_n, args.cpu_passes, threads, smoothers, spawn_coeff,
                                            cores=set(avail[:threads]))
        serial_value, serial_s, sm_mesh = cpu_baseline(args.cpu_serial_n, 1, 1, smoothers, spawn_coeff,
                                                       cores={avail[0]})
        model = cpu_model()
        cpu = {"value": cpu_value, "unit": "smoothed vertices/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle (glibc libm), Thread mode on {threads} pinned cores: {args.cpu_passes} "
                         f"pass(es) each of {','.join(smoothers)} on the {args.cpu_n}^3 workload mesh "
                         f"({cm.nverts} vertices), {cpu_s:.1f} s",
               "cpu_model": model,
               "serial": {"value": serial_value, "cores": 1,
                          "sample": f"Serial mode on one pinned core: 1 pass each of {','.join(smoothers)} on the "
                                    f"{args.cpu_serial_n}^3 construction ({sm_mesh.nverts} vertices), {serial_s:.1f} s"}}
        log(f"cpu oracle ({model}): {cpu_value:.0f} vertices/s on {threads} threads ({cpu_s:.1f} s); serial "
            f"{serial_value:.0f} vertices/s ({serial_s:.1f} s)")
        if topology is not None:
            # The oracle's restructure of the topology leg's first mesh, one core.
            from gpumesh_amd import cube_mesh
            from oracle.pyoracle import OracleMesh
            tm = cube_mesh(args.topo_n, "tet", 0.35, seed=2)
            om = OracleMesh.from_cube(tm)
            om.set_crew("mean_ratio", "uniform", scaling=float(args.topo_n))
            t0 = time.perf_counter()
            oc = om.restructure(5, 5, 0.5, 1.5)
            t_or = time.perf_counter() - t0
            if oc != topology["passes"][0]["counts"]:
                raise RuntimeError("topology leg: the first restructure's operation counts differ from the oracle's")
            cpu["topology"] = {"restructure_s": t_or, "tets_per_s": len(tm.tets) / t_or, "cores": 1,
                               "sample": f"oracle BATR restructure of the {args.topo_n}^3 tet cube "
                                         f"({len(tm.tets)} tets), one core"}
            log(f"cpu oracle restructure: {t_or:.2f} s")

    ms_per_step = total_wall / (args.steps * len(smoothers)) * 1e3
    out = {
        "metric": "smoothed vertices/s on 1M-vertex tet mesh; p50/p99 latency batched stream",
        "value": value,
        "unit": "vertices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32_polish": "f32+f64", "fp64": "f64", "fp64_fast": "f64", "fp32": "f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"Kuhn tet cube {args.n}^3 cells ({m.nverts} vertices, {gm.ntets} tets), "
                               f"jitter 0.35h seed 2, mean ratio, {'+'.join(smoothers)} x {args.steps} passes, "
                               f"{args.precision}, {args.kernel} kernel",
                   "vertices": m.nverts, "movable": gm.nmovable, "tets": gm.ntets, "colour_groups": gm.ngroups,
                   "smoothers": smoothers, "passes": args.steps, "precision": args.precision,
                   "kernel": args.kernel, "spawn_coeff": spawn_coeff},
        "per_smoother": {sm: {"vertices_per_s": gm.nmovable * args.steps / p["wall_s"],
                              "device_ms": p["res"]["graph_total_ms"],
                              "eager_device_ms": p["res"]["total_ms"],
                              "final_min": float(p["res"]["min"][-1]), "final_mean": float(p["res"]["mean"][-1]),
                              "elements_evaluated_per_s": (float(np.sum(p["res"]["eval_counts"])) *
                                                           gm.sum_patch / gm.nmovable +
                                                           (args.steps + 1) * gm.ntets) / p["wall_s"]}
                         for sm, p in per.items()},
        "plain_fp64_vertices_per_s": plain_value,
        "speedup_vs_plain_fp64": (value / plain_value) if plain_value else None,
        "quality_vs_fp64": quality,
        "headline_within_1e-6_of_fp64": headline_equal,
        "fp64_exact_vertices_per_s": exact_value,
        "equal_quality": equal_quality,
        "batched": batched,
        "serving": serving,
        "topology": topology,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
