"""Python handle of a device-resident mesh (a thin layer over the C ABI).

Names follow the reference's smoothing vocabulary: smoothers
(Nelder-Mead, gradient descent, spawn search, GETMe), evaluators (mean ratio,
metric conformity), samplers (uniform, analytic, texture), colour groups
(NodeGroups' independent sets) and passes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Schedule, check, lib, ptr

SMOOTHERS = {"nm": 0, "nelder_mead": 0, "gd": 1, "gradient_descent": 1, "spawn": 2, "spawn_search": 2,
             "getme": 3}
EVALUATORS = {"mean_ratio": 0, "metric_conformity": 1}
SAMPLERS = {"uniform": 0, "analytic": 1, "texture": 2, "analytic_reference": 3}
PRECISIONS = {"fp64": 0, "fp32": 1, "fp32_polish": 2, "fp64_fast": 3}
KERNELS = {"plain": 0, "wave": 1}
# gm_kernel_class order (include/gpumesh.h).
KERNEL_CLASSES = ["smooth_f32", "smooth_f64", "getme_elem_f32", "getme_elem_f64", "getme_vert_f32",
                  "getme_vert_f64", "evaluate"]
KERNEL_CLASSES_N = len(KERNEL_CLASSES)


class GpuMesh:
    """A mesh uploaded to the current HIP device (or, with device=False, its
    host-side topology only).  Vertex ids are the renumbered ones."""

    def __init__(self, pos, fixed, tets, hexs, device: bool = True, pris=None):
        """pris: prisms, 6 vertices each (MeshPri order), or None."""
        L = lib()
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        fixed = np.ascontiguousarray(fixed, np.int32).reshape(-1)
        tets = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
        hexs = np.ascontiguousarray(hexs, np.uint32).reshape(-1, 8)
        if pris is not None and len(pris):
            self._create_desc(pos, fixed, tets, hexs, None, [], None, device, pris)
            return
        h = C.c_void_p()
        fn = L.gm_mesh_create if device else L.gm_mesh_create_host
        check(fn(pos.shape[0], ptr(pos), ptr(fixed), tets.shape[0], ptr(tets), hexs.shape[0], ptr(hexs),
                 C.byref(h)), "gm_mesh_create" if device else "gm_mesh_create_host")
        self._h = h
        self.device = device
        self._read_info()

    def _read_info(self):
        L = lib()
        info = np.zeros(8, np.int64)
        check(L.gm_mesh_info(self._h, ptr(info)), "gm_mesh_info")
        (self.nverts, self.ntets, self.nhexs, self.nmovable, self.ngroups, self.sum_patch,
         self.sum_nbr, self.max_patch) = (int(x) for x in info)
        n = C.c_int64(0)
        check(L.gm_mesh_prisms(self._h, C.byref(n), None), "gm_mesh_prisms")
        self.npris = int(n.value)

    def _create_desc(self, pos, fixed, tets, hexs, constraint, constraints, groups, device, pris):
        L = lib()
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        fixed = np.ascontiguousarray(fixed, np.int32).reshape(-1)
        tets = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
        hexs = np.ascontiguousarray(hexs, np.uint32).reshape(-1, 8)
        pris = np.ascontiguousarray(np.zeros((0, 6)) if pris is None else pris, np.uint32).reshape(-1, 6)
        table = (_lib.Constraint * max(len(constraints), 1))()
        for k, (kind, a, b) in enumerate(constraints):
            table[k].kind = int(kind)
            for i in range(3):
                table[k].a[i] = float(a[i])
                table[k].b[i] = float(b[i])
        keep = [pos, fixed, tets, hexs, pris]
        d = _lib.MeshDesc()
        d.nverts, d.pos, d.fixed = pos.shape[0], ptr(pos), ptr(fixed)
        d.ntets, d.tets, d.nhexs, d.hexs = tets.shape[0], ptr(tets), hexs.shape[0], ptr(hexs)
        d.npris, d.pris = pris.shape[0], ptr(pris)
        if constraint is not None:
            cons = np.ascontiguousarray(constraint, np.int32).reshape(-1)
            keep.append(cons)
            d.constraint, d.nconstraints, d.constraints = ptr(cons), len(constraints), C.addressof(table)
        d.ngroups = -1
        if groups is not None:
            off = np.ascontiguousarray(groups[0], np.int64)
            nodes = np.ascontiguousarray(groups[1], np.uint32)
            keep += [off, nodes]
            d.ngroups, d.group_off, d.group_nodes = off.size - 1, ptr(off), ptr(nodes)
        d.device_copy = int(device)
        h = C.c_void_p()
        check(L.gm_mesh_create_desc(C.byref(d), C.byref(h)), "gm_mesh_create_desc")
        del keep
        self._h = h
        self.device = device
        self._read_info()

    @classmethod
    def constrained(cls, pos, fixed, tets, hexs, constraint, constraints, groups=None, device: bool = True,
                    pris=None):
        """A mesh whose vertices may slide on planes or segments
        (gm_mesh_create_desc): constraint[v] indexes `constraints` (rows
        (kind, a, b), kind 1 plane through a with normal b, 2 segment a-b) or
        is -1; groups = (group_off, group_nodes) keeps the caller's numbering
        and colouring, None colours and renumbers; pris as in the constructor."""
        self = cls.__new__(cls)
        self._create_desc(pos, fixed, tets, hexs, constraint, constraints, groups, device, pris)
        return self

    @classmethod
    def grouped(cls, pos, fixed, tets, hexs, group_off, group_nodes, device: bool = True):
        """A mesh in the caller's numbering with the caller's colour groups
        (gm_mesh_create_grouped): no renumbering, groups checked on the host."""
        L = lib()
        self = cls.__new__(cls)
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        fixed = np.ascontiguousarray(fixed, np.int32).reshape(-1)
        tets = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
        hexs = np.ascontiguousarray(hexs, np.uint32).reshape(-1, 8)
        off = np.ascontiguousarray(group_off, np.int64)
        nodes = np.ascontiguousarray(group_nodes, np.uint32)
        h = C.c_void_p()
        check(L.gm_mesh_create_grouped(pos.shape[0], ptr(pos), ptr(fixed), tets.shape[0], ptr(tets), hexs.shape[0],
                                       ptr(hexs), off.size - 1, ptr(off), ptr(nodes), int(device), C.byref(h)),
              "gm_mesh_create_grouped")
        self._h = h
        self.device = device
        self._read_info()
        return self

    def device_bytes(self) -> int:
        """Device memory held by the handle."""
        b = np.zeros(1, np.int64)
        check(lib().gm_mesh_device_bytes(self._h, ptr(b)), "gm_mesh_device_bytes")
        return int(b[0])

    @classmethod
    def from_cube(cls, m, device: bool = True):
        return cls(m.pos, m.fixed, m.tets, m.hexs, device=device, pris=getattr(m, "pris", None))

    def prisms(self) -> np.ndarray:
        """Renumbered prisms (P, 6)."""
        out = np.zeros((max(self.npris, 1), 6), np.uint32)
        check(lib().gm_mesh_prisms(self._h, None, ptr(out)), "gm_mesh_prisms")
        return out[: self.npris]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gm_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----- topology
    def permutation(self) -> np.ndarray:
        out = np.empty(self.nverts, np.int64)
        check(lib().gm_mesh_permutation(self._h, ptr(out)), "gm_mesh_permutation")
        return out

    def groups(self):
        off = np.empty(self.ngroups + 1, np.int64)
        nodes = np.empty(max(self.nmovable, 1), np.uint32)
        check(lib().gm_mesh_groups(self._h, ptr(off), ptr(nodes)), "gm_mesh_groups")
        return off, nodes[: self.nmovable]

    def elements(self):
        t = np.empty(max(4 * self.ntets, 1), np.uint32)
        h = np.empty(max(8 * self.nhexs, 1), np.uint32)
        check(lib().gm_mesh_elements(self._h, ptr(t), ptr(h)), "gm_mesh_elements")
        return t[: 4 * self.ntets].reshape(-1, 4), h[: 8 * self.nhexs].reshape(-1, 8)

    def topology(self):
        eo = np.empty(self.nverts + 1, np.uint32)
        er = np.empty(max(self.sum_patch, 1), np.uint32)
        no = np.empty(self.nverts + 1, np.uint32)
        nb = np.empty(max(self.sum_nbr, 1), np.uint32)
        check(lib().gm_mesh_topology(self._h, ptr(eo), ptr(er), ptr(no), ptr(nb)), "gm_mesh_topology")
        return eo, er[: self.sum_patch], no, nb[: self.sum_nbr]

    # ----- positions
    def positions(self) -> np.ndarray:
        out = np.empty(3 * self.nverts, np.float64)
        check(lib().gm_mesh_get_positions(self._h, ptr(out)), "gm_mesh_get_positions")
        return out.reshape(-1, 3)

    def set_positions(self, pos) -> None:
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
        check(lib().gm_mesh_set_positions(self._h, ptr(pos)), "gm_mesh_set_positions")

    # ----- crew
    def set_aspect_ratio(self, aspect: float) -> None:
        """Aspect ratio of the reference's analytic field ("analytic_reference")."""
        check(lib().gm_mesh_set_aspect_ratio(self._h, float(aspect)), "gm_mesh_set_aspect_ratio")

    def set_crew(self, evaluator="mean_ratio", sampler="uniform", scaling=1.0, texture=None) -> None:
        args = [None, None, None, None]
        if sampler == "texture":
            size, tmin, text, texels = texture
            keep = [np.ascontiguousarray(size, np.int32), np.ascontiguousarray(tmin, np.float64),
                    np.ascontiguousarray(text, np.float64), np.ascontiguousarray(texels, np.float32)]
            args = [ptr(a) for a in keep]
        check(lib().gm_mesh_set_crew(self._h, EVALUATORS[evaluator], SAMPLERS[sampler], float(scaling), *args),
              "gm_mesh_set_crew")

    # ----- evaluation and smoothing
    def evaluate(self, element_qualities: bool = False):
        out = np.zeros(3, np.float64)
        hist = np.zeros(20, np.int32)
        ne = self.ntets + self.npris + self.nhexs
        q = np.zeros(max(ne, 1), np.float64) if element_qualities else None
        check(lib().gm_mesh_evaluate(self._h, ptr(out), ptr(hist), ptr(q)), "gm_mesh_evaluate")
        r = {"min": float(out[0]), "inv_sum": float(out[1]), "n": int(out[2]),
             "mean": float(out[2] / out[1]) if out[1] != 0 else float("nan"), "hist": hist}
        if element_qualities:
            r["elem_qual"] = q[:ne]
        return r

    def smooth_groups(self, smoother: str, group_begin: int, group_end: int, precision: str = "fp64",
                      kernel: str = "wave", spawn_coeff: float = 1.0 / 20.0) -> np.ndarray:
        """One partial sweep over colour groups [group_begin, group_end); returns
        the per-vertex patch-evaluation counts."""
        cnt = np.zeros(self.nverts, np.int32)
        check(lib().gm_mesh_smooth_groups(self._h, SMOOTHERS[smoother], int(group_begin), int(group_end),
                                          PRECISIONS[precision], KERNELS[kernel], float(spawn_coeff), ptr(cnt)),
              "gm_mesh_smooth_groups")
        return cnt

    def smooth(self, smoother: str, passes: int, precision: str = "fp64", kernel: str = "plain",
               spawn_coeff: float = 1.0 / 20.0, counts: bool = True):
        stats = np.zeros((passes + 1, 3), np.float64)
        hist = np.zeros((passes + 1, 20), np.int32)
        cnt = np.zeros(self.nverts, np.int32) if counts else None
        check(lib().gm_mesh_smooth(self._h, SMOOTHERS[smoother], int(passes), PRECISIONS[precision],
                                   KERNELS[kernel], float(spawn_coeff), ptr(stats), ptr(hist), ptr(cnt)),
              "gm_mesh_smooth")
        t = np.zeros(2, np.float64)
        check(lib().gm_mesh_last_timing(self._h, ptr(t)), "gm_mesh_last_timing")
        kms = np.zeros(KERNEL_CLASSES_N, np.float64)
        kn = np.zeros(KERNEL_CLASSES_N, np.int64)
        check(lib().gm_mesh_kernel_times(self._h, ptr(kms), ptr(kn)), "gm_mesh_kernel_times")
        r = {"min": stats[:, 0].copy(), "mean": stats[:, 1].copy(), "inv_sum": stats[:, 2].copy(),
             "hist": hist, "total_ms": float(t[0]), "sweep_ms": float(t[1]),
             "kernel_ms": dict(zip(KERNEL_CLASSES, kms.tolist())),
             "kernel_launches": dict(zip(KERNEL_CLASSES, kn.tolist()))}
        if counts:
            r["eval_counts"] = cnt
        return r

    def restructure(self, topo_passes: int = 5, refinement_sweeps: int = 5, min_edge: float = 0.5,
                    max_edge: float = 1.5) -> dict:
        """BATR topology operations (gm_mesh_restructure): face swaps, edge
        swaps and edge splits/merges in the crew's metric, then the mesh is
        rebuilt (new vertex and tet counts, numbering and colour groups).
        Returns the operation counts."""
        counts = np.zeros(5, np.int64)
        check(lib().gm_mesh_restructure(self._h, int(topo_passes), int(refinement_sweeps), float(min_edge),
                                        float(max_edge), ptr(counts)), "gm_mesh_restructure")
        self._read_info()
        return dict(zip(("face_swaps", "edge_swaps", "splits", "merges", "passes"), (int(c) for c in counts)))

    def smooth_global(self, smoother: str, global_passes: int, relocation_passes: int, topo_passes: int = 5,
                      refinement_sweeps: int = 5, min_edge: float = 0.5, max_edge: float = 1.5,
                      precision: str = "fp64", kernel: str = "wave", spawn_coeff: float = 1.0 / 20.0) -> list:
        """The reference's global loop with topology operations on
        (AbstractVertexWiseSmoother.cpp:53-80): per global pass, restructure,
        then `relocation_passes` sweeps on the device; returns each global
        pass's operation counts and pass records."""
        out = []
        for _ in range(int(global_passes)):
            counts = self.restructure(topo_passes, refinement_sweeps, min_edge, max_edge)
            r = self.smooth(smoother, relocation_passes, precision=precision, kernel=kernel, spawn_coeff=spawn_coeff)
            out.append({"counts": counts, **r})
        return out

    def smooth_autopilot(self, smoother: str, max_passes: int, min_qual_threshold: float = 0.001,
                         qual_mean_threshold: float = 0.0, precision: str = "fp64", kernel: str = "wave",
                         spawn_coeff: float = 1.0 / 20.0, counts: bool = True):
        """Smoothing under the reference's autopilot (gm_mesh_smooth_schedule):
        sweeps continue while the harmonic-mean gain exceeds min_qual_threshold
        or the minimum gain exceeds qual_mean_threshold, at most max_passes.
        Returns the records of the passes run ("passes" = sweeps run)."""
        stats = np.zeros((max_passes + 1, 3), np.float64)
        hist = np.zeros((max_passes + 1, 20), np.int32)
        cnt = np.zeros(self.nverts, np.int32) if counts else None
        sched = Schedule(1, float(min_qual_threshold), float(qual_mean_threshold))
        run = np.zeros(1, np.int32)
        check(lib().gm_mesh_smooth_schedule(self._h, SMOOTHERS[smoother], int(max_passes), C.byref(sched),
                                            PRECISIONS[precision], KERNELS[kernel], float(spawn_coeff), ptr(stats),
                                            ptr(hist), ptr(cnt), ptr(run)), "gm_mesh_smooth_schedule")
        k = int(run[0]) + 1
        r = {"passes": int(run[0]), "min": stats[:k, 0].copy(), "mean": stats[:k, 1].copy(),
             "inv_sum": stats[:k, 2].copy(), "hist": hist[:k].copy()}
        if counts:
            r["eval_counts"] = cnt
        return r


def set_topology_cache(max_bytes: int) -> None:
    """gm_set_topology_cache: meshes of identical connectivity share one host
    topology (and, while it is cached, its device arrays); 0 turns it off."""
    check(lib().gm_set_topology_cache(int(max_bytes)), "gm_set_topology_cache")


def topology_cache_stats() -> dict:
    out = np.zeros(4, np.int64)
    check(lib().gm_topology_cache_stats(ptr(out)), "gm_topology_cache_stats")
    return {"hits": int(out[0]), "misses": int(out[1]), "entries": int(out[2]), "bytes": int(out[3])}


def prepare_topology(pos, fixed, tets, hexs, device: bool = True) -> None:
    """gm_topology_prepare: builds a mesh shape's topology into the cache
    (and its device arrays when `device`) without creating a mesh."""
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    fixed = np.ascontiguousarray(fixed, np.int32).reshape(-1)
    tets = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
    hexs = np.ascontiguousarray(hexs, np.uint32).reshape(-1, 8)
    check(lib().gm_topology_prepare(pos.shape[0], ptr(pos), ptr(fixed), tets.shape[0], ptr(tets), hexs.shape[0],
                                    ptr(hexs), int(bool(device))), "gm_topology_prepare")


def _timings(L, fn_t, fn_k, h):
    t = np.zeros(2, np.float64)
    check(fn_t(h, ptr(t)), "last_timing")
    kms = np.zeros(KERNEL_CLASSES_N, np.float64)
    kn = np.zeros(KERNEL_CLASSES_N, np.int64)
    check(fn_k(h, ptr(kms), ptr(kn)), "kernel_times")
    return {"total_ms": float(t[0]), "sweep_ms": float(t[1]),
            "kernel_ms": dict(zip(KERNEL_CLASSES, kms.tolist())),
            "kernel_launches": dict(zip(KERNEL_CLASSES, kn.tolist()))}


class GpuBatch:
    """Several resident meshes smoothed by fused launches (one launch per
    colour step and element kind over all members).  The members stay owned
    by their GpuMesh objects, which must outlive the batch."""

    def __init__(self, meshes):
        self.meshes = list(meshes)
        arr = (C.c_void_p * len(self.meshes))(*[m._h.value for m in self.meshes])
        h = C.c_void_p()
        check(lib().gm_batch_create(arr, len(self.meshes), C.byref(h)), "gm_batch_create")
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, smoother: str, patches_per_group: int, tile_elems: int) -> None:
        check(lib().gm_batch_set_shape(self._h, SMOOTHERS[smoother], int(patches_per_group), int(tile_elems)),
              "gm_batch_set_shape")

    def autotune(self, smoother: str, precision: str = "fp32", passes: int = 1):
        """Times every launch shape; keeps the fastest.  Returns ((ppg, tile), times[4, 3] ms)."""
        best = np.zeros(2, np.int32)
        times = np.zeros(12, np.float64)
        check(lib().gm_batch_autotune(self._h, SMOOTHERS[smoother], PRECISIONS[precision], int(passes), ptr(best),
                                      ptr(times)), "gm_batch_autotune")
        return (int(best[0]), int(best[1])), times.reshape(4, 3)

    def smooth(self, smoother: str, passes: int, precision: str = "fp64", kernel: str = "wave",
               spawn_coeff: float = 1.0 / 20.0, graph: bool = False):
        n = len(self.meshes)
        stats = np.zeros((n, passes + 1, 3), np.float64)
        hist = np.zeros((n, passes + 1, 20), np.int32)
        L = lib()
        check(L.gm_batch_smooth(self._h, SMOOTHERS[smoother], int(passes), PRECISIONS[precision], KERNELS[kernel],
                                float(spawn_coeff), int(bool(graph)), ptr(stats), ptr(hist)), "gm_batch_smooth")
        r = {"min": stats[:, :, 0].copy(), "mean": stats[:, :, 1].copy(), "inv_sum": stats[:, :, 2].copy(),
             "hist": hist}
        r.update(_timings(L, L.gm_batch_last_timing, L.gm_batch_kernel_times, self._h))
        return r

    def smooth_autopilot(self, smoother: str, max_passes: int, min_qual_threshold: float = 0.001,
                         qual_mean_threshold: float = 0.0, precision: str = "fp64", kernel: str = "wave",
                         spawn_coeff: float = 1.0 / 20.0, graph: bool = False):
        """GpuMesh.smooth_autopilot over every member (each stops on its own
        records); "passes" holds the sweeps each member ran, records past a
        member's last one are zero."""
        n = len(self.meshes)
        stats = np.zeros((n, max_passes + 1, 3), np.float64)
        hist = np.zeros((n, max_passes + 1, 20), np.int32)
        run = np.zeros(n, np.int32)
        sched = Schedule(1, float(min_qual_threshold), float(qual_mean_threshold))
        L = lib()
        check(L.gm_batch_smooth_schedule(self._h, SMOOTHERS[smoother], int(max_passes), C.byref(sched),
                                         PRECISIONS[precision], KERNELS[kernel], float(spawn_coeff),
                                         int(bool(graph)), ptr(stats), ptr(hist), ptr(run)),
              "gm_batch_smooth_schedule")
        r = {"passes": run, "min": stats[:, :, 0].copy(), "mean": stats[:, :, 1].copy(),
             "inv_sum": stats[:, :, 2].copy(), "hist": hist}
        r.update(_timings(L, L.gm_batch_last_timing, L.gm_batch_kernel_times, self._h))
        return r
