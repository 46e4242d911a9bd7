"""Equal quality: the FP64 modes against the FP64 oracle at the parity
tolerances of this path (positions within 1e-9 h, final min and mean within
1e-12).

* GM_FP64_FAST runs every pass on the mixed-precision FP64 kernels (FP64
  arithmetic in another operation order, d^(-2/3) from a hardware seed
  refined to FP64); its trajectories stay on the exact ones.
* The exact FP64 kernels use the project's correctly rounded pow / cbrt /
  exp; against the oracle calling glibc's (the reference's libm) they differ
  only in last bits, within the same tolerances.
"""
import numpy as np
import pytest

from gpumesh_amd import GpuMesh, cube_mesh

pytestmark = pytest.mark.gpu

POS_TOL_H = 1e-9
Q_TOL = 1e-12
CASES = [("tet", 20, "nm"), ("tet", 20, "gd"), ("tet", 20, "spawn"), ("tet", 20, "getme"),
         ("hex", 12, "nm"), ("hex", 12, "gd"), ("hex", 12, "spawn"), ("hex", 12, "getme")]


def _run(oracle, kind, n, smoother, precision, libm):
    m = cube_mesh(n, kind, 0.3, seed=1)
    coeff = 0.25 if smoother == "spawn" else 1.0 / 20.0
    oracle.set_libm(libm)
    try:
        om = oracle.OracleMesh.from_cube(m)
        ref = om.smooth(smoother, 10, threads=8, spawn_coeff=coeff)
    finally:
        oracle.set_libm(1)
    gm = GpuMesh.from_cube(m)
    out = gm.smooth(smoother, 10, precision=precision, kernel="wave", spawn_coeff=coeff)
    dx = float(np.max(np.abs(gm.positions() - om.positions()))) * n
    return dx, out, ref


@pytest.mark.parametrize("kind,n,smoother", CASES)
def test_fp64_fast_equal_quality(oracle, kind, n, smoother):
    dx, out, ref = _run(oracle, kind, n, smoother, "fp64_fast", 1)
    print(f"fp64_fast {kind} {smoother}: max |dx|/h {dx:.2e}, dmin {out['min'][-1] - ref['min'][-1]:+.2e}, "
          f"dmean {out['mean'][-1] - ref['mean'][-1]:+.2e}")
    assert dx <= POS_TOL_H
    assert np.max(np.abs(out["min"] - ref["min"])) <= Q_TOL
    assert np.max(np.abs(out["mean"] - ref["mean"])) <= Q_TOL


@pytest.mark.parametrize("kind,n,smoother", CASES)
def test_exact_fp64_against_glibc_oracle(oracle, kind, n, smoother):
    dx, out, ref = _run(oracle, kind, n, smoother, "fp64", 0)
    print(f"fp64 vs glibc oracle {kind} {smoother}: max |dx|/h {dx:.2e}, "
          f"dmin {out['min'][-1] - ref['min'][-1]:+.2e}, dmean {out['mean'][-1] - ref['mean'][-1]:+.2e}")
    assert dx <= POS_TOL_H
    assert np.max(np.abs(out["min"] - ref["min"])) <= Q_TOL
    assert np.max(np.abs(out["mean"] - ref["mean"])) <= Q_TOL
