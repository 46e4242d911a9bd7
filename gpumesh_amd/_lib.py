"""ctypes loader of libgpumesh_hip.so (the HIP smoothing library).

The library is built in-tree (gpumesh_amd/libgpumesh_hip.so) by
`__graft_entry__.build()` / `make -C gpumesh_amd/csrc`.  There is no fallback:
if the library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpumesh_hip.so")
_LIB = None

GM_OK = 0
STATUS = {0: "GM_OK", 1: "GM_ERR_ARG", 2: "GM_ERR_HIP", 3: "GM_ERR_UNSUPPORTED", 4: "GM_ERR_NOMEM"}

# Every symbol the C ABI exports (include/gpumesh.h).
EXPORTS = [
    "gm_last_error", "gm_version", "gm_device_count", "gm_mesh_create", "gm_mesh_create_host",
    "gm_mesh_destroy", "gm_mesh_info", "gm_mesh_permutation", "gm_mesh_groups", "gm_mesh_elements",
    "gm_mesh_topology", "gm_mesh_get_positions", "gm_mesh_set_positions", "gm_mesh_set_crew",
    "gm_mesh_evaluate", "gm_mesh_smooth", "gm_mesh_smooth_groups", "gm_mesh_last_timing", "gm_mesh_kernel_times",
    "gm_mesh_device_bytes", "gm_batch_create", "gm_batch_destroy", "gm_batch_set_shape", "gm_batch_smooth",
    "gm_batch_last_timing", "gm_batch_kernel_times", "gm_batch_autotune", "gm_server_create", "gm_server_destroy", "gm_server_set_shape", "gm_server_submit", "gm_server_wait",
    "gm_server_result", "gm_server_release", "gm_server_stats", "gm_cube_mesh", "gm_set_topology_cache", "gm_set_device_pool",
    "gm_topology_cache_stats", "gm_math_probe", "gm_mesh_smooth_schedule", "gm_batch_smooth_schedule",
    "gm_job_size", "gm_job_encode", "gm_job_info", "gm_job_decode", "gm_server_submit_job",
    "gm_set_polish_passes", "gm_evaluate_noalloc", "gm_mesh_create_grouped", "gm_mesh_set_aspect_ratio",
    "gm_mesh_create_desc", "gm_topology_prepare", "gm_throughput_probe", "gm_server_set_texture", "gm_mesh_prisms",
    "gm_mesh_restructure",
]


class Constraint(C.Structure):
    """gm_constraint: a plane (a = point, b = normal) or a segment (a, b ends)."""
    _fields_ = [("kind", C.c_int), ("a", C.c_double * 3), ("b", C.c_double * 3)]


class MeshDesc(C.Structure):
    """gm_mesh_desc."""
    _fields_ = [("nverts", C.c_int64), ("pos", C.c_void_p), ("fixed", C.c_void_p), ("ntets", C.c_int64),
                ("tets", C.c_void_p), ("nhexs", C.c_int64), ("hexs", C.c_void_p), ("constraint", C.c_void_p),
                ("nconstraints", C.c_int64), ("constraints", C.c_void_p), ("ngroups", C.c_int64),
                ("group_off", C.c_void_p), ("group_nodes", C.c_void_p), ("device_copy", C.c_int),
                ("npris", C.c_int64), ("pris", C.c_void_p)]


class Schedule(C.Structure):
    """gm_schedule: autopilot flag and the two gain thresholds (reference names)."""
    _fields_ = [("autopilot", C.c_int), ("min_qual_threshold", C.c_double), ("qual_mean_threshold", C.c_double)]


class GpuMeshError(RuntimeError):
    """A C-ABI call returned an error status."""

    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: {STATUS.get(status, status)}: {message}")
        self.status = status


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C gpumesh_amd/csrc")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.gm_last_error.restype = C.c_char_p
    L.gm_version.restype = C.c_char_p
    L.gm_device_count.argtypes = [C.POINTER(C.c_int)]
    for name in ("gm_mesh_create", "gm_mesh_create_host"):
        getattr(L, name).argtypes = [i64, vp, vp, i64, vp, i64, vp, C.POINTER(vp)]
    L.gm_mesh_create_grouped.argtypes = [i64, vp, vp, i64, vp, i64, vp, i64, vp, vp, C.c_int, C.POINTER(vp)]
    L.gm_mesh_create_desc.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
    L.gm_mesh_destroy.argtypes = [vp]
    L.gm_mesh_info.argtypes = [vp, vp]
    L.gm_mesh_permutation.argtypes = [vp, vp]
    L.gm_mesh_groups.argtypes = [vp, vp, vp]
    L.gm_mesh_elements.argtypes = [vp, vp, vp]
    L.gm_mesh_topology.argtypes = [vp, vp, vp, vp, vp]
    L.gm_mesh_get_positions.argtypes = [vp, vp]
    L.gm_mesh_set_positions.argtypes = [vp, vp]
    L.gm_mesh_set_crew.argtypes = [vp, C.c_int, C.c_int, dbl, vp, vp, vp, vp]
    L.gm_mesh_evaluate.argtypes = [vp, vp, vp, vp]
    L.gm_mesh_set_aspect_ratio.argtypes = [vp, dbl]
    L.gm_mesh_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dbl, vp, vp, vp]
    L.gm_mesh_smooth_groups.argtypes = [vp, C.c_int, i64, i64, C.c_int, C.c_int, dbl, vp]
    L.gm_mesh_last_timing.argtypes = [vp, vp]
    L.gm_mesh_kernel_times.argtypes = [vp, vp, vp]
    L.gm_mesh_device_bytes.argtypes = [vp, vp]
    L.gm_batch_create.argtypes = [vp, i64, C.POINTER(vp)]
    L.gm_batch_destroy.argtypes = [vp]
    L.gm_batch_set_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.gm_batch_smooth.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, dbl, C.c_int, vp, vp]
    L.gm_job_size.argtypes = [i64, i64, i64]
    L.gm_job_size.restype = i64
    L.gm_job_encode.argtypes = [i64, vp, vp, i64, vp, i64, vp, C.c_int, C.c_int, C.c_int, vp, i64]
    L.gm_job_info.argtypes = [vp, i64, vp, vp]
    L.gm_job_decode.argtypes = [vp, i64, vp, vp, vp, vp]
    L.gm_server_submit_job.argtypes = [vp, vp, i64, C.POINTER(i64)]
    L.gm_mesh_smooth_schedule.argtypes = [vp, C.c_int, C.c_int, C.POINTER(Schedule), C.c_int, C.c_int, dbl, vp, vp,
                                          vp, vp]
    L.gm_batch_smooth_schedule.argtypes = [vp, C.c_int, C.c_int, C.POINTER(Schedule), C.c_int, C.c_int, dbl,
                                           C.c_int, vp, vp, vp]
    L.gm_batch_last_timing.argtypes = [vp, vp]
    L.gm_batch_kernel_times.argtypes = [vp, vp, vp]
    L.gm_batch_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp]
    L.gm_server_create.argtypes = [i64, C.c_int, C.c_int, C.c_int, C.c_int, dbl, C.POINTER(vp)]
    L.gm_server_destroy.argtypes = [vp]
    L.gm_server_set_shape.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.gm_server_set_texture.argtypes = [vp, vp, vp, vp, vp]
    L.gm_mesh_prisms.argtypes = [vp, vp, vp]
    L.gm_mesh_restructure.argtypes = [vp, C.c_int, C.c_int, dbl, dbl, vp]
    L.gm_server_submit.argtypes = [vp, i64, vp, vp, i64, vp, i64, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(i64)]
    L.gm_server_wait.argtypes = [vp, i64, dbl, C.POINTER(C.c_int)]
    L.gm_server_result.argtypes = [vp, i64, vp, vp, vp, vp]
    L.gm_server_release.argtypes = [vp, i64]
    L.gm_server_stats.argtypes = [vp, vp]
    L.gm_set_topology_cache.argtypes = [i64]
    L.gm_set_device_pool.argtypes = [i64]
    L.gm_set_polish_passes.argtypes = [C.c_int]
    L.gm_evaluate_noalloc.argtypes = [vp, i64, vp, i64, vp, i64, C.c_int, C.c_int, dbl, vp, vp, vp]
    L.gm_topology_cache_stats.argtypes = [vp]
    L.gm_topology_prepare.argtypes = [i64, vp, vp, i64, vp, i64, vp, C.c_int]
    L.gm_throughput_probe.argtypes = [C.c_int, vp]
    L.gm_cube_mesh.argtypes = [C.c_int, C.c_int, dbl, C.c_uint64, dbl, vp, vp, vp]
    L.gm_math_probe.argtypes = [C.c_int, vp, vp, vp, i64]
    for name in EXPORTS:
        if name not in ("gm_last_error", "gm_version", "gm_job_size"):
            getattr(L, name).restype = C.c_int
    _LIB = L
    return L


def check(status: int, where: str) -> None:
    if status != GM_OK:
        raise GpuMeshError(status, where, lib().gm_last_error().decode())


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def device_count() -> int:
    n = C.c_int(0)
    check(lib().gm_device_count(C.byref(n)), "gm_device_count")
    return int(n.value)


def math_probe(op: int, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float64)
    b = None if b is None else np.ascontiguousarray(b, np.float64)
    out = np.empty_like(a)
    check(lib().gm_math_probe(int(op), ptr(a), ptr(b), ptr(out), a.size), "gm_math_probe")
    return out
