"""Writes tests/golden/c1_golden.json: fixtures of the plain-method workload
(20^3 Kuhn cube, jitter 0.3 h, seed 1, mean ratio, Nelder-Mead, 10 passes)
computed by the FP64 oracle.  They pin the oracle (and, through the GPU
parity tests, the product) against regressions: colouring/permutation
digests, per-pass min/mean with both libm modes, and a digest of the final
positions.  Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpumesh_amd import cube_mesh  # noqa: E402
from oracle import pyoracle  # noqa: E402


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    m = cube_mesh(20, "tet", 0.3, seed=1)
    out = {"workload": "Kuhn cube 20^3, jitter 0.3h, seed 1, mean ratio, metric-free, Nelder-Mead x 10",
           "nverts": m.nverts, "ntets": int(m.tets.shape[0]), "input_positions_sha256": digest(m.pos)}
    for mode in (0, 1):
        pyoracle.set_libm(mode)
        om = pyoracle.OracleMesh.from_cube(m)
        if mode == 1:
            groups, tr = om.groups()
            out["permutation_sha256"] = digest(om.permutation())
            out["group_interior_ranges"] = groups[:, 4:6].tolist()
            t, _ = om.elements()
            out["renumbered_tets_sha256"] = digest(t)
        r = om.smooth("nm", 10, threads=8)
        key = "glibc" if mode == 0 else "gm_libm"
        out[key] = {"min": r["min"].tolist(), "mean": r["mean"].tolist(),
                    "hist_final": r["hist"][-1].tolist(),
                    "eval_count_total": int(r["eval_counts"].sum()),
                    "final_positions_sha256": digest(om.positions())}
    pyoracle.set_libm(1)
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
