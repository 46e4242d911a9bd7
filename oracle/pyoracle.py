"""ctypes binding of the FP64 CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
as the checker.  The product (gpumesh_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SMOOTHERS = {"nm": 0, "gd": 1, "spawn": 2, "getme": 3}
EVALUATORS = {"mean_ratio": 0, "metric_conformity": 1}
SAMPLERS = {"uniform": 0, "analytic": 1, "texture": 2, "analytic_reference": 3}

_P = np.ctypeslib.ndpointer


def build() -> str:
    """Compiles liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    vp = C.c_void_p
    i64 = C.c_int64
    L.orc_mesh_create.restype = vp
    L.orc_mesh_create.argtypes = [i64, _P(np.float64), _P(np.int32), i64, _P(np.uint32), i64, _P(np.uint32)]
    L.orc_mesh_destroy.argtypes = [vp]
    L.orc_mesh_create_constrained.restype = vp
    L.orc_mesh_create_constrained.argtypes = [i64, _P(np.float64), _P(np.int32), _P(np.int32), i64, _P(np.int32),
                                              _P(np.float64), i64, _P(np.uint32), i64, _P(np.uint32)]
    L.orc_mesh_create_full.restype = vp
    L.orc_mesh_create_full.argtypes = [i64, _P(np.float64), _P(np.int32), vp, i64, vp, vp, i64, _P(np.uint32), i64,
                                       _P(np.uint32), i64, _P(np.uint32)]
    L.orc_mesh_npris.argtypes = [vp]
    L.orc_mesh_npris.restype = i64
    L.orc_mesh_prisms.argtypes = [vp, _P(np.uint32)]
    L.orc_constraint_snap.argtypes = [C.c_int, _P(np.float64), _P(np.float64), _P(np.float64), _P(np.float64)]
    L.orc_set_libm.argtypes = [C.c_int]
    L.orc_set_libm.restype = C.c_int
    L.orc_mesh_nverts.argtypes = [vp]
    L.orc_mesh_nverts.restype = i64
    L.orc_mesh_positions.argtypes = [vp, _P(np.float64)]
    L.orc_mesh_set_positions.argtypes = [vp, _P(np.float64)]
    L.orc_mesh_permutation.argtypes = [vp, _P(np.int64)]
    L.orc_mesh_elements.argtypes = [vp, _P(np.uint32), _P(np.uint32)]
    L.orc_mesh_topology_sizes.argtypes = [vp, _P(np.int64)]
    L.orc_mesh_topology.argtypes = [vp, _P(np.int64), _P(np.uint64), _P(np.int64), _P(np.uint32)]
    L.orc_mesh_ngroups.argtypes = [vp]
    L.orc_mesh_ngroups.restype = i64
    L.orc_mesh_groups.argtypes = [vp, _P(np.int64), _P(np.int64)]
    L.orc_set_crew.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, vp, vp, vp, vp]
    L.orc_set_aspect.argtypes = [vp, C.c_double]
    L.orc_evaluate.argtypes = [vp, C.c_int, _P(np.float64), _P(np.int32)]
    for name in ("orc_tet_quality", "orc_hex_quality", "orc_tet_volume", "orc_hex_volume", "orc_pri_quality",
                 "orc_pri_volume"):
        getattr(L, name).argtypes = [vp, _P(np.float64)]
        getattr(L, name).restype = C.c_double
    L.orc_element_qualities.argtypes = [vp, _P(np.float64)]
    L.orc_gm_libm.argtypes = [C.c_int, _P(np.float64), _P(np.float64), i64]
    L.orc_gm_libm2.argtypes = [C.c_int, _P(np.float64), _P(np.float64), _P(np.float64), i64]
    L.orc_patch_quality.argtypes = [vp, C.c_uint32]
    L.orc_patch_quality.restype = C.c_double
    L.orc_local_size.argtypes = [vp, C.c_uint32]
    L.orc_local_size.restype = C.c_double
    L.orc_metric_at.argtypes = [vp, _P(np.float64), _P(np.float64)]
    L.orc_riemannian_distance.argtypes = [vp, _P(np.float64), _P(np.float64)]
    L.orc_riemannian_distance.restype = C.c_double
    L.orc_trace_vertex.argtypes = [vp, C.c_int, C.c_double, C.c_uint32, _P(np.float64), i64]
    L.orc_trace_vertex.restype = i64
    L.orc_getme_element.argtypes = [vp, C.c_int, _P(np.float64), _P(np.float64), _P(np.float64)]
    L.orc_specified_metric.argtypes = [vp, _P(np.float64), _P(np.float64)]
    L.orc_smooth_vertices.argtypes = [vp, C.c_int, C.c_double, _P(np.uint32), i64, vp]
    L.orc_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, vp, vp, vp]
    L.orc_smooth.restype = C.c_int
    L.orc_smooth_schedule.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                      C.c_int, C.c_double, vp, vp, vp, vp]
    L.orc_smooth_schedule.restype = C.c_int
    L.orc_mesh_ntets.argtypes = [vp]
    L.orc_mesh_ntets.restype = i64
    L.orc_mesh_fixed.argtypes = [vp, _P(np.int32)]
    L.orc_mesh_restructure.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, _P(np.int64)]
    L.orc_mesh_restructure.restype = C.c_int
    _LIB = L
    return L


def gm_libm(op: int, x, y=None) -> np.ndarray:
    """The project's libm evaluated on the host (op 2 pow23, 3 cbrt, 4 exp,
    6 cos, 7 log, 8 pow(x, y))."""
    x = np.ascontiguousarray(x, np.float64)
    y = x if y is None else np.ascontiguousarray(y, np.float64)
    out = np.empty_like(x)
    lib().orc_gm_libm2(int(op), x, y, out, x.size)
    return out


def constraint_snap(kind: int, a, b, pos) -> np.ndarray:
    """One snap of `pos` onto a plane (kind 1: through a, normal b) or a
    segment (kind 2: from a to b)."""
    out = np.empty(3, np.float64)
    lib().orc_constraint_snap(int(kind), np.ascontiguousarray(a, np.float64), np.ascontiguousarray(b, np.float64),
                              np.ascontiguousarray(pos, np.float64), out)
    return out


def set_libm(mode: int) -> int:
    """0: glibc std::pow/std::exp (the reference's calls); 1: the project's gm_libm."""
    return lib().orc_set_libm(int(mode))


class OracleMesh:
    """An oracle mesh: topology + node groups built (and renumbered) at creation."""

    def __init__(self, pos, fixed, tets, hexs, constraint=None, constraints=None, pris=None):
        """constraint[v]: index into `constraints` or -1; constraints: rows
        (kind, a, b) with kind 1 = plane (a point, b normal), 2 = segment.
        pris: prisms, 6 vertices each (MeshPri)."""
        L = lib()
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        fixed = np.ascontiguousarray(fixed, dtype=np.int32).reshape(-1)
        tets = np.ascontiguousarray(tets, dtype=np.uint32).reshape(-1, 4)
        hexs = np.ascontiguousarray(hexs, dtype=np.uint32).reshape(-1, 8)
        pris = np.ascontiguousarray(np.zeros((0, 6)) if pris is None else pris, dtype=np.uint32).reshape(-1, 6)
        self.ntet = tets.shape[0]
        self.nhex = hexs.shape[0]
        self.npri = pris.shape[0]
        self.nverts = pos.shape[0]
        if self.npri:
            one = np.zeros(1, np.uint32)
            keep = []
            cptr = kptr = pptr = None
            ncons = 0
            if constraint is not None:
                cons = np.ascontiguousarray(constraint, np.int32).reshape(-1)
                kinds = np.ascontiguousarray([int(c[0]) for c in constraints] or [0], np.int32)
                params = np.ascontiguousarray([list(c[1]) + list(c[2]) for c in constraints] or [[0.0] * 6],
                                              np.float64).reshape(-1)
                keep = [cons, kinds, params]
                cptr, kptr, pptr = (a.ctypes.data for a in keep)
                ncons = len(constraints)
            self._h = L.orc_mesh_create_full(self.nverts, pos.reshape(-1), fixed, cptr, ncons, kptr, pptr, self.ntet,
                                             tets.reshape(-1) if self.ntet else one, self.npri, pris.reshape(-1),
                                             self.nhex, hexs.reshape(-1) if self.nhex else one)
        elif constraint is None:
            self._h = L.orc_mesh_create(self.nverts, pos.reshape(-1), fixed, self.ntet, tets.reshape(-1),
                                        self.nhex, hexs.reshape(-1))
        else:
            cons = np.ascontiguousarray(constraint, np.int32).reshape(-1)
            kinds = np.ascontiguousarray([int(c[0]) for c in constraints] or [0], np.int32)
            params = np.ascontiguousarray([list(c[1]) + list(c[2]) for c in constraints] or [[0.0] * 6],
                                          np.float64).reshape(-1)
            self._h = L.orc_mesh_create_constrained(self.nverts, pos.reshape(-1), fixed, cons, len(constraints),
                                                    kinds, params, self.ntet, tets.reshape(-1), self.nhex,
                                                    hexs.reshape(-1))
        self.evaluator = "mean_ratio"

    @classmethod
    def from_cube(cls, m):
        return cls(m.pos, m.fixed, m.tets, m.hexs, pris=getattr(m, "pris", None))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _LIB is not None:
            _LIB.orc_mesh_destroy(h)
            self._h = None

    # --- state
    def positions(self) -> np.ndarray:
        out = np.empty(3 * self.nverts, np.float64)
        lib().orc_mesh_positions(self._h, out)
        return out.reshape(-1, 3)

    def set_positions(self, pos) -> None:
        lib().orc_mesh_set_positions(self._h, np.ascontiguousarray(pos, np.float64).reshape(-1))

    def permutation(self) -> np.ndarray:
        out = np.empty(self.nverts, np.int64)
        lib().orc_mesh_permutation(self._h, out)
        return out

    def elements(self):
        t = np.empty(4 * self.ntet, np.uint32)
        h = np.empty(8 * self.nhex, np.uint32)
        lib().orc_mesh_elements(self._h, t if t.size else np.empty(1, np.uint32), h if h.size else np.empty(1, np.uint32))
        return t.reshape(-1, 4), h.reshape(-1, 8)

    def topology(self):
        L = lib()
        sz = np.empty(2, np.int64)
        L.orc_mesh_topology_sizes(self._h, sz)
        eo = np.empty(self.nverts + 1, np.int64)
        er = np.empty(max(int(sz[0]), 1), np.uint64)
        no = np.empty(self.nverts + 1, np.int64)
        nb = np.empty(max(int(sz[1]), 1), np.uint32)
        L.orc_mesh_topology(self._h, eo, er, no, nb)
        return eo, er[: int(sz[0])], no, nb[: int(sz[1])]

    def groups(self):
        L = lib()
        g = int(L.orc_mesh_ngroups(self._h))
        r = np.empty(max(6 * g, 1), np.int64)
        t = np.empty(8, np.int64)
        L.orc_mesh_groups(self._h, r, t)
        return r[: 6 * g].reshape(-1, 6), t.reshape(4, 2)

    # --- crew
    def set_crew(self, evaluator="mean_ratio", sampler="uniform", measurer=None, scaling=1.0,
                 texture=None):
        L = lib()
        meas = -1 if measurer is None else {"metric_free": 0, "metric_wise": 1}[measurer]
        args = [None, None, None, None]
        keep = []
        if sampler == "texture":
            size, tmin, text, texels = texture
            size = np.ascontiguousarray(size, np.int32)
            tmin = np.ascontiguousarray(tmin, np.float64)
            text = np.ascontiguousarray(text, np.float64)
            texels = np.ascontiguousarray(texels, np.float32)
            keep = [size, tmin, text, texels]
            args = [a.ctypes.data for a in keep]
        L.orc_set_crew(self._h, EVALUATORS[evaluator], SAMPLERS[sampler], meas, float(scaling), *args)
        self.evaluator = evaluator

    def set_aspect_ratio(self, aspect: float) -> None:
        lib().orc_set_aspect(self._h, float(aspect))

    # --- queries
    def evaluate(self, chunks: int = 1):
        out = np.empty(3, np.float64)
        hist = np.empty(20, np.int32)
        lib().orc_evaluate(self._h, int(chunks), out, hist)
        return {"min": float(out[0]), "inv_sum": float(out[1]), "n": int(out[2]),
                "mean": float(out[2] / out[1]), "hist": hist}

    def prisms(self) -> np.ndarray:
        p = np.empty(max(6 * self.npri, 1), np.uint32)
        lib().orc_mesh_prisms(self._h, p)
        return p[: 6 * self.npri].reshape(-1, 6)

    def element_qualities(self) -> np.ndarray:
        """Tets, then prisms, then hexes."""
        n = self.ntet + self.npri + self.nhex
        out = np.empty(max(n, 1), np.float64)
        lib().orc_element_qualities(self._h, out)
        return out[:n]

    def tet_quality(self, vp) -> float:
        return lib().orc_tet_quality(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1))

    def hex_quality(self, vp) -> float:
        return lib().orc_hex_quality(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1))

    def tet_volume(self, vp) -> float:
        return lib().orc_tet_volume(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1))

    def hex_volume(self, vp) -> float:
        return lib().orc_hex_volume(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1))

    def pri_quality(self, vp) -> float:
        return lib().orc_pri_quality(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1))

    def pri_volume(self, vp) -> float:
        return lib().orc_pri_volume(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1))

    def patch_quality(self, v: int) -> float:
        return lib().orc_patch_quality(self._h, int(v))

    def local_size(self, v: int) -> float:
        return lib().orc_local_size(self._h, int(v))

    def metric_at(self, p) -> np.ndarray:
        out = np.empty(9, np.float64)
        lib().orc_metric_at(self._h, np.ascontiguousarray(p, np.float64), out)
        return out.reshape(3, 3).T   # rows/cols of the symmetric metric

    def riemannian_distance(self, a, b) -> float:
        return lib().orc_riemannian_distance(self._h, np.ascontiguousarray(a, np.float64),
                                             np.ascontiguousarray(b, np.float64))

    def trace_vertex(self, smoother: str, v: int, spawn_coeff: float = 1.0 / 20.0) -> np.ndarray:
        """Smooths vertex v alone and returns every patch evaluation it made,
        in order, as rows (x, y, z, quality)."""
        cap = 4096
        out = np.empty(4 * cap, np.float64)
        n = int(lib().orc_trace_vertex(self._h, SMOOTHERS[smoother], float(spawn_coeff), int(v), out, cap))
        return out[: 4 * min(n, cap)].reshape(-1, 4)

    def getme_element(self, vp):
        """GETMe element transform of free-standing vertices: (vpp, weight)."""
        vp = np.ascontiguousarray(vp, np.float64).reshape(-1, 3)
        out = np.empty(vp.size, np.float64)
        w = np.empty(1, np.float64)
        kind = {4: 0, 6: 2, 8: 1}[vp.shape[0]]
        lib().orc_getme_element(self._h, kind, vp.reshape(-1), out, w)
        return out.reshape(-1, 3), float(w[0])

    def specified_metric(self, vp) -> np.ndarray:
        """Keast-rule metric of a tetrahedron (MetricConformityEvaluator::specifiedMetric)."""
        out = np.empty(9, np.float64)
        lib().orc_specified_metric(self._h, np.ascontiguousarray(vp, np.float64).reshape(-1), out)
        return out.reshape(3, 3).T

    def smooth_vertices(self, smoother: str, vids, spawn_coeff: float = 1.0 / 20.0):
        vids = np.ascontiguousarray(vids, np.uint32)
        cnt = np.zeros(vids.size, np.int64)
        lib().orc_smooth_vertices(self._h, SMOOTHERS[smoother], float(spawn_coeff), vids, vids.size,
                                  cnt.ctypes.data)
        return cnt

    def smooth(self, smoother: str, passes: int, threads: int = 1, order: int = 1, eval_chunks: int = 1,
               spawn_coeff: float = 1.0 / 20.0):
        """Runs `passes` sweeps; returns per-pass stats, histograms, per-vertex eval counts."""
        stats = np.zeros((passes + 1, 3), np.float64)
        hist = np.zeros((passes + 1, 20), np.int32)
        cnt = np.zeros(self.nverts, np.int64)
        rc = lib().orc_smooth(self._h, SMOOTHERS[smoother], int(passes), int(threads), int(order),
                              int(eval_chunks), float(spawn_coeff), stats.ctypes.data, hist.ctypes.data,
                              cnt.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"oracle smooth failed: {rc}")
        return {"min": stats[:, 0].copy(), "mean": stats[:, 1].copy(), "inv_sum": stats[:, 2].copy(),
                "hist": hist, "eval_counts": cnt}

    def smooth_autopilot(self, smoother: str, max_passes: int, min_qual_threshold: float = 0.001,
                         qual_mean_threshold: float = 0.0, threads: int = 1, order: int = 1, eval_chunks: int = 1,
                         spawn_coeff: float = 1.0 / 20.0):
        """The autopilot schedule (Schedule's thresholds, applied as AbstractSmoother does), at most
        `max_passes` sweeps; records of the sweeps run plus the initial one."""
        stats = np.zeros((max_passes + 1, 3), np.float64)
        hist = np.zeros((max_passes + 1, 20), np.int32)
        cnt = np.zeros(self.nverts, np.int64)
        run = C.c_int(0)
        rc = lib().orc_smooth_schedule(self._h, SMOOTHERS[smoother], int(max_passes), 1, float(min_qual_threshold),
                                       float(qual_mean_threshold), int(threads), int(order), int(eval_chunks),
                                       float(spawn_coeff), stats.ctypes.data, hist.ctypes.data, cnt.ctypes.data,
                                       C.addressof(run))
        if rc != 0:
            raise RuntimeError(f"oracle smooth failed: {rc}")
        k = run.value + 1
        return {"passes": run.value, "min": stats[:k, 0].copy(), "mean": stats[:k, 1].copy(),
                "inv_sum": stats[:k, 2].copy(), "hist": hist[:k].copy(), "eval_counts": cnt}

    # --- topology operations
    def fixed(self) -> np.ndarray:
        out = np.empty(self.nverts, np.int32)
        lib().orc_mesh_fixed(self._h, out)
        return out

    def restructure(self, topo_passes: int = 5, refinement_sweeps: int = 5, min_edge: float = 0.5,
                    max_edge: float = 1.5, libm: int = 0) -> dict:
        """BATR topology operations (face swaps, edge swaps, edge splits and
        merges) on a tetrahedral mesh, then the mesh is recompiled and
        renumbered; returns the operation counts.  libm: 0 = glibc (the
        reference's, and the library's host topologist's), 1 = the project's."""
        counts = np.zeros(5, np.int64)
        old = set_libm(libm)
        try:
            rc = lib().orc_mesh_restructure(self._h, int(topo_passes), int(refinement_sweeps), float(min_edge),
                                            float(max_edge), counts)
        finally:
            set_libm(old)
        if rc != 0:
            raise RuntimeError(f"oracle restructure refused: {rc}")
        self.nverts = int(lib().orc_mesh_nverts(self._h))
        self.ntet = int(lib().orc_mesh_ntets(self._h))
        return dict(zip(("face_swaps", "edge_swaps", "splits", "merges", "passes"), (int(c) for c in counts)))

    def smooth_global(self, smoother: str, global_passes: int, relocation_passes: int, topo_passes: int = 5,
                      refinement_sweeps: int = 5, min_edge: float = 0.5, max_edge: float = 1.5, threads: int = 1,
                      spawn_coeff: float = 1.0 / 20.0) -> list:
        """The reference's global loop with topology operations on
        (AbstractVertexWiseSmoother.cpp:53-80): per global pass, restructure
        (glibc), then `relocation_passes` sweeps (the current libm mode);
        returns each pass's counts and records."""
        out = []
        for _ in range(global_passes):
            counts = self.restructure(topo_passes, refinement_sweeps, min_edge, max_edge)
            r = self.smooth(smoother, relocation_passes, threads=threads, spawn_coeff=spawn_coeff)
            out.append({"counts": counts, **r})
        return out
