"""gpumesh_amd -- MI355X-native vertex-relocation mesh smoothing.

Hand-written HIP kernels for gfx950 behind a C ABI (include/gpumesh.h,
libgpumesh_hip.so); this package is the Python binding, the synthetic
cube-mesh generator used by the tests and the benchmark, and mesh files (the
reference's JSON layout, binary job files).
"""
from ._lib import GpuMeshError, device_count, lib, math_probe  # noqa: F401
from .mesh import GpuBatch, GpuMesh, prepare_topology, set_topology_cache, topology_cache_stats  # noqa: F401
from .meshgen import CubeMesh, box_boundary, cube_mesh  # noqa: F401
from .meshio import MeshData, decode_job, encode_job, read_json, write_json  # noqa: F401
from .serving import GpuServer  # noqa: F401

__all__ = ["GpuBatch", "GpuMesh", "GpuServer", "GpuMeshError", "CubeMesh", "box_boundary", "cube_mesh", "device_count", "lib",
           "math_probe", "MeshData", "read_json", "write_json", "encode_job", "decode_job", "prepare_topology",
           "set_topology_cache", "topology_cache_stats"]
