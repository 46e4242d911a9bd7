"""gm_evaluate_noalloc (the allocation-free evaluation) against
gm_mesh_evaluate on the same meshes, through the C ABI: two calls in a row,
the second on a smaller mesh of another element kind, bit-identical minimum
and histogram and the same sum of 1/q (bound 1e-12 relative; the reduction
tree is the same, so it is expected to be equal).  Device buffers come from
the HIP runtime the library links (the test allocates; the entry point does
not)."""
import ctypes as C

import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh
from gpumesh_amd._lib import GpuMeshError, check, lib

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        lib()                                   # loads libamdhip64 through the library first
        _HIP = C.CDLL("libamdhip64.so.7")
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipFree.argtypes = [C.c_void_p]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipDeviceSynchronize.argtypes = []
    return _HIP


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(int(nbytes), 8)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, 1) == 0    # host to device

    def get(self, dtype, n):
        out = np.empty(n, dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        hip().hipFree(self.p)


def noalloc(pos, tets, hexs, evaluator=0, sampler=0, scaling=1.0):
    bufs = [DevBuf(pos.nbytes), DevBuf(tets.nbytes), DevBuf(hexs.nbytes), DevBuf(24), DevBuf(80)]
    try:
        bufs[0].put(pos)
        bufs[1].put(tets)
        bufs[2].put(hexs)
        check(lib().gm_evaluate_noalloc(bufs[0].p, pos.shape[0], bufs[1].p, tets.shape[0], bufs[2].p,
                                        hexs.shape[0], evaluator, sampler, float(scaling), bufs[3].p, bufs[4].p,
                                        None), "gm_evaluate_noalloc")
        assert hip().hipDeviceSynchronize() == 0
        return bufs[3].get(np.float64, 3), bufs[4].get(np.int32, 20)
    finally:
        for b in bufs:
            b.free()


def _compare(m, evaluator="mean_ratio", sampler="uniform", scaling=1.0):
    gm = GpuMesh.from_cube(m)
    gm.set_crew(evaluator, sampler, scaling=scaling)
    ref = gm.evaluate()
    out, hist = noalloc(np.ascontiguousarray(m.pos), np.ascontiguousarray(m.tets, np.uint32),
                        np.ascontiguousarray(m.hexs, np.uint32), {"mean_ratio": 0, "metric_conformity": 1}[evaluator],
                        {"uniform": 0, "analytic": 1}[sampler], scaling)
    assert out[0] == ref["min"]
    assert np.array_equal(hist, ref["hist"])
    assert abs(out[1] - ref["inv_sum"]) <= 1e-12 * abs(ref["inv_sum"])
    assert out[2] == ref["n"]
    return out[1] == ref["inv_sum"]


def test_noalloc_matches_mesh_evaluate_twice():
    big = cube_mesh(24, "tet", 0.3, seed=1)          # 82,944 tets, 81 evaluation blocks
    small = cube_mesh(6, "hex", 0.3, seed=2)         # 216 hexes, 1 block
    assert _compare(big)
    assert _compare(small)
    # The anisotropic crew on the small mesh, after the larger call.
    assert _compare(small, "metric_conformity", "analytic", scaling=6.0)


def test_noalloc_empty_mesh_and_refused_sizes():
    out, hist = noalloc(np.zeros((0, 3)), np.zeros((0, 4), np.uint32), np.zeros((0, 8), np.uint32))
    assert out.tolist() == [1.0, 0.0, 0.0] and not hist.any()
    with pytest.raises(GpuMeshError):
        check(lib().gm_evaluate_noalloc(None, 4, None, 2 ** 31 + 1, None, 0, 0, 0, 1.0, None, None, None),
              "gm_evaluate_noalloc")
    with pytest.raises(GpuMeshError):
        check(lib().gm_evaluate_noalloc(None, 4, None, 1, None, 0, 0, 2, 1.0, None, None, None),
              "gm_evaluate_noalloc")
