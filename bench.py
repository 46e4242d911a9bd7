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
give identical results, supplies the per-launch kernel times.  Also reported: the plain FP64 single-mesh kernel (one lane
per vertex) on the same workload, the roofline of the dominant kernel and the
CPU oracle on a bounded sample of the same work.

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
# corner terms: determinant 6, squared norm 8, |d|^(-2/3) and 3/q 4,
# accumulation 2.
FLOP_AFFINE_TET = 20.0
# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (each scaled by the same counter's reading for a
# 1 GiB device copy: x2.0 and x1.0) on the 100^3 Kuhn cube.
MEASURED_TRAFFIC = {("spawnMfmaKernel<0, 4>", 100): 109167442.0}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_flop(res, smoother, mesh):
    """FLOP of the smoothing kernels of one run, from counted patch evaluations."""
    patch = mesh.sum_patch / max(mesh.nmovable, 1)
    evals = float(np.sum(res["eval_counts"]))
    flop = evals * patch * FLOP_MR_TET_PATCH
    if smoother == "getme":
        passes = len(res["min"]) - 1
        return flop, passes * mesh.ntets * FLOP_GETME_TET
    return flop, 0.0


def cpu_baseline(n, passes, threads, smoothers, spawn_coeff):
    """The FP64 CPU oracle (Thread mode, `threads` workers) on a bounded sample."""
    from gpumesh_amd import cube_mesh
    from oracle import pyoracle
    pyoracle.set_libm(0)
    m = cube_mesh(n, "tet", 0.35, seed=2)
    moved = 0
    t_total = 0.0
    for sm in smoothers:
        om = pyoracle.OracleMesh.from_cube(m)
        t0 = time.perf_counter()
        om.smooth(sm, passes, threads=threads, order=1, eval_chunks=threads, spawn_coeff=spawn_coeff)
        t_total += time.perf_counter() - t0
        moved += m.nmovable * passes
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
    reset()
    t0 = time.perf_counter()
    b.smooth("getme", passes, precision="fp32", graph=False)
    wall_eager = time.perf_counter() - t0
    # The same meshes one by one (FP32, then the plain FP64 kernel).
    reset()
    t0 = time.perf_counter()
    for g in gms:
        g.smooth("getme", passes, precision="fp32", kernel="wave", counts=False)
    wall_single = time.perf_counter() - t0
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
            "final_min_mean": [float(np.min(r["min"][:, -1])), float(np.mean(r["mean"][:, -1]))]}


def serving_leg(args):
    """Poisson job stream through the serving loop at several rates."""
    from gpumesh_amd import GpuMesh, GpuServer, cube_mesh
    rng = np.random.default_rng(5)
    jobs = []
    for k in range(args.serve_jobs):
        n = int(round(np.exp(rng.uniform(np.log(8.0), np.log(100.0)))))
        kind = "tet" if rng.random() < 0.5 else "hex"
        jitter = float(rng.uniform(0.2, 0.4))
        passes = int(rng.integers(5, 21))
        smoother = ["nm", "gd", "getme"][int(rng.integers(0, 3))]
        jobs.append((n, kind, jitter, 5000 + k, smoother, passes))
    t0 = time.perf_counter()
    meshes = [cube_mesh(n, kind, jit, seed=sd) for n, kind, jit, sd, _, _ in jobs]
    log(f"serving: {len(jobs)} jobs, {sum(m.nverts for m in meshes)} vertices in total, generated in "
        f"{time.perf_counter() - t0:.2f}s")
    work = sum(m.nmovable * j[5] for m, j in zip(meshes, jobs))
    rates = [float(x) for x in args.serve_rates.split(",") if x]
    out = {"workload": f"{len(jobs)} jobs per rate: tet/hex 50/50 cubes, n log-uniform [8,100], jitter U[0.2,0.4]h, "
                       f"5-20 passes, NM/GD/GETMe uniform, fp32 + fp64 polish, Poisson arrivals",
           "jobs": len(jobs), "vertex_passes_per_run": work, "rates": {}}
    with GpuServer(budget_bytes=288 * 10**9, worker_threads=args.serve_workers) as srv:
        # Untimed warm-up jobs (first launches of each kernel).
        warm = [srv.submit_mesh(meshes[0], sm, 1) for sm in ("nm", "gd", "getme")]
        for w in warm:
            srv.wait(w)
            srv.release(w)
        for rate in rates:
            arr = np.cumsum(np.random.default_rng(int(rate) + 5).exponential(1.0 / rate, len(jobs)))
            ids = []
            t0 = time.perf_counter()
            for (n, kind, jit, sd, sm, p), m, at in zip(jobs, meshes, arr):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids.append(srv.submit_mesh(m, sm, p, "fp32_polish"))
            lat = []
            for j in ids:
                st = srv.wait(j, 120000)
                if st != "done":
                    raise RuntimeError(f"serving job ended {st}")
                lat.append(srv.result(j)["latency_ms"])
            span = time.perf_counter() - t0
            lat = np.array(lat)
            out["rates"][str(int(rate))] = {
                "offered_jobs_per_s": rate, "jobs_per_s": len(jobs) / span, "vertices_per_s": work / span,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "max_ms": float(lat.max()), "span_s": span}
            log(f"serving rate {rate:.0f}/s: {len(jobs) / span:.1f} jobs/s, {work / span / 1e6:.1f} M vertices/s, "
                f"p50 {np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms")
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
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed passes per smoother")
    ap.add_argument("--warmup", type=int, default=2, help="untimed passes per smoother before timing")
    ap.add_argument("--n", type=int, default=100, help="cells per cube edge")
    ap.add_argument("--smoothers", default="gd,getme,spawn")
    ap.add_argument("--precision", default="fp32_polish", choices=["fp64", "fp32", "fp32_polish"])
    ap.add_argument("--kernel", default="wave", choices=["plain", "wave"])
    ap.add_argument("--plain-steps", type=int, default=2, help="passes of the plain FP64 kernel per smoother")
    ap.add_argument("--cpu-n", type=int, default=80, help="cube size of the CPU sample")
    ap.add_argument("--cpu-passes", type=int, default=2, help="passes per smoother of the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-meshes", type=int, default=256)
    ap.add_argument("--batch-passes", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=48)
    ap.add_argument("--serve-rates", default="50,200,1000,2000")
    ap.add_argument("--serve-plain-sample", type=int, default=8)
    ap.add_argument("--serve-workers", type=int, default=8, help="host worker threads of the server")
    ap.add_argument("--no-serve", action="store_true")
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
    from gpumesh_amd._lib import check, lib
    check(lib().gm_set_topology_cache(8 << 30), "gm_set_topology_cache")
    if args.serve_only:
        print(json.dumps(serving_leg(args)))
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

    # Dominant kernel class of the timed runs, and its roofline.
    best = None
    for sm, p in per.items():
        res = p["res"]
        f_smooth, f_getme = algorithmic_flop(res, sm, gm)
        km, kn = res["kernel_ms"], res["kernel_launches"]
        classes = {
            "smooth_f32": (km["smooth_f32"], kn["smooth_f32"], None),
            "getme_elem_f32": (km["getme_elem_f32"], kn["getme_elem_f32"], None),
            "getme_vert_f32": (km["getme_vert_f32"], kn["getme_vert_f32"], None),
            "evaluate": (km["evaluate"], kn["evaluate"], None),
        }
        for cls, (ms, launches, _) in classes.items():
            if launches == 0:
                continue
            if best is None or ms > best["ms"]:
                best = {"smoother": sm, "class": cls, "ms": ms, "launches": launches, "res": res,
                        "f_smooth": f_smooth, "f_getme": f_getme}
    roofline = None
    if best is not None:
        avg_s = best["ms"] / best["launches"] / 1e3
        steps = args.steps
        f32_passes = steps - 1 if args.precision == "fp32_polish" else steps
        if best["class"] == "smooth_f32":
            # FP32 passes only (the last pass is the FP64 polish): scale the
            # counted FLOP to the FP32 share of the passes.
            flop_per_launch = best["f_smooth"] * (f32_passes / steps) / best["launches"]
            achieved = flop_per_launch / avg_s / 1e12
            kname = {"spawn": "spawnMfmaKernel<0, 4>", "gd": "gdMfmaTetKernel<3>",
                     "nm": "nmQuadKernel<6>"}.get(best["smoother"], "smoothWaveKernel<float>")
            executed = flop_per_launch * FLOP_AFFINE_TET / FLOP_MR_TET_PATCH / avg_s / 1e12
            # The spawn-search and gradient-descent kernels put the corner
            # terms' linear parts on the matrix cores; the FP32 peak is the
            # same for the matrix cores and packed vector FP32.
            roofline = {"bound": "mfma" if best["smoother"] in ("spawn", "gd") else "valu", "achieved": achieved,
                        "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS,
                        "traffic": MEASURED_TRAFFIC.get((kname, args.n)), "kernel": kname,
                        "avg_launch_ms": avg_s * 1e3, "flop_per_launch": flop_per_launch,
                        "executed_tflops": executed, "executed_frac_of_nonpacked_fp32": executed / (PEAK_FP32_TFLOPS / 2)}
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
                flop_per_launch = (2 * FLOP_MR_TET_PATCH + FLOP_GETME_PULL) * gm.sum_patch / best["launches"] * f32_passes
                kname = "getmeWaveKernel<float, 0, 0, 32, 0>"
            achieved = flop_per_launch / avg_s / 1e12
            roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": kname,
                        "avg_launch_ms": avg_s * 1e3, "flop_per_launch": flop_per_launch}

    # Final quality of the mixed-precision runs against FP64 runs of the same
    # passes (the FP64 wave kernels are bit-identical to the plain FP64 ones).
    quality = {}
    for sm, p in per.items():
        gm.set_positions(start)
        r64 = gm.smooth(sm, args.steps, precision="fp64", kernel="wave", spawn_coeff=spawn_coeff, counts=False)
        quality[sm] = {"fp64_final_min": float(r64["min"][-1]), "fp64_final_mean": float(r64["mean"][-1]),
                       "min_diff": float(p["res"]["min"][-1] - r64["min"][-1]),
                       "mean_diff": float(p["res"]["mean"][-1] - r64["mean"][-1])}
        log(f"{sm}: final min/mean vs fp64: {quality[sm]['min_diff']:+.2e} / {quality[sm]['mean_diff']:+.2e}")

    batched = None if args.no_batch else batched_leg(args)
    serving = None if args.no_serve else serving_leg(args)

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(threads, 16))
        cpu_value, cpu_s, cm = cpu_baseline(args.cpu_n, args.cpu_passes, threads, smoothers, spawn_coeff)
        cpu = {"value": cpu_value, "unit": "smoothed vertices/s", "cores": threads, "kind": "port",
               "sample": f"FP64 oracle, Thread mode, {args.cpu_passes} passes each of {','.join(smoothers)} on the same "
                         f"construction at {args.cpu_n}^3 cells ({cm.nverts} vertices), glibc libm; "
                         f"{cpu_s:.1f} s"}
        log(f"cpu oracle: {cpu_value:.0f} vertices/s on {threads} threads ({cpu_s:.1f} s)")

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
        "dtype": "f32+f64" if args.precision == "fp32_polish" else ("f64" if args.precision == "fp64" else "f32"),
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
        "batched": batched,
        "serving": serving,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
