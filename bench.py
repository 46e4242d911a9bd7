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
timed region).  Also reported: the plain FP64 single-mesh kernel (one lane
per vertex) on the same workload, the roofline of the dominant kernel and the
CPU oracle on a bounded sample of the same work.

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

# Algorithmic FLOP, counting +, -, *, /, sqrt and pow as one each (the
# derivation is in DESIGN.md).  One mean-ratio tetrahedron evaluated inside a
# patch: 3 metric-free edge vectors 18, F = E Fr^-1 45, det 14, |F|^2 17,
# corner quality 4 (incl. pow), patch accumulation 3 (1/q and two adds).
FLOP_MR_TET_PATCH = 101.0
# One GETMe tetrahedron transform: 4 face normals 60, 4 scaled moves 56,
# 2 volumes 50, volume ratio and cube root 3, centre 12, rescale 36,
# 2 qualities 196, weight 2.
FLOP_GETME_TET = 415.0
# GETMe vertex update, per patch element pulled: 3 mul + 4 add.
FLOP_GETME_PULL = 7.0


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
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: multi-GPU runs are not part of this build "
              "(meshes shard with no exchange; see the sharding design notes).", file=sys.stderr)
        sys.exit(2)

    import gpumesh_amd
    from gpumesh_amd import GpuMesh, cube_mesh

    if gpumesh_amd.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    smoothers = [s for s in args.smoothers.split(",") if s]
    spawn_coeff = 0.25

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
    for sm in smoothers:
        if args.warmup > 0:
            gm.set_positions(start)
            gm.smooth(sm, args.warmup, precision=args.precision, kernel=args.kernel, spawn_coeff=spawn_coeff)
        gm.set_positions(start)
        t0 = time.perf_counter()
        res = gm.smooth(sm, args.steps, precision=args.precision, kernel=args.kernel, spawn_coeff=spawn_coeff)
        wall = time.perf_counter() - t0
        per[sm] = {"wall_s": wall, "res": res}
        log(f"{sm}: {args.steps} passes {wall * 1e3:.1f} ms wall ({res['total_ms']:.1f} ms device), "
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
        if best["class"] == "smooth_f32":
            # FP32 passes only (the last pass is the FP64 polish): scale the
            # counted FLOP to the FP32 share of the passes.
            f32_passes = steps - 1 if args.precision == "fp32_polish" else steps
            flop_per_launch = best["f_smooth"] * (f32_passes / steps) / best["launches"]
            achieved = flop_per_launch / avg_s / 1e12
            roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                        "kernel": f"{'spawnLaneKernel' if best['smoother'] == 'spawn' else ('gdLaneKernel' if best['smoother'] == 'gd' else 'smoothWaveKernel')}<float> ({best['smoother']})", "avg_launch_ms": avg_s * 1e3,
                        "flop_per_launch": flop_per_launch}
        elif best["class"] == "evaluate":
            bytes_per = 16.0 * gm.ntets + 32.0 * gm.nverts
            achieved = bytes_per / avg_s / 1e9
            roofline = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": None, "kernel": "evalKernel",
                        "avg_launch_ms": avg_s * 1e3, "bytes_per_launch": bytes_per}
        else:
            f32_passes = steps - 1 if args.precision == "fp32_polish" else steps
            if best["class"] == "getme_elem_f32":
                flop_per_launch = FLOP_GETME_TET * gm.ntets
            else:
                flop_per_launch = (2 * FLOP_MR_TET_PATCH + FLOP_GETME_PULL) * gm.sum_patch / best["launches"] * f32_passes
            achieved = flop_per_launch / avg_s / 1e12
            roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": best["class"],
                        "avg_launch_ms": avg_s * 1e3, "flop_per_launch": flop_per_launch}

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
        "metric": "smoothed vertices/s on 1M-vertex tet mesh",
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
                              "device_ms": p["res"]["total_ms"],
                              "final_min": float(p["res"]["min"][-1]), "final_mean": float(p["res"]["mean"][-1]),
                              "elements_evaluated_per_s": (float(np.sum(p["res"]["eval_counts"])) *
                                                           gm.sum_patch / gm.nmovable +
                                                           (args.steps + 1) * gm.ntets) / p["wall_s"]}
                         for sm, p in per.items()},
        "plain_fp64_vertices_per_s": plain_value,
        "speedup_vs_plain_fp64": (value / plain_value) if plain_value else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
