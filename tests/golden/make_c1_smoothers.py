"""Writes tests/golden/c1_smoothers.json: fixtures of the plain-method mesh
(20^3 Kuhn cube, jitter 0.3 h, seed 1, mean ratio, metric-free) computed by
the FP64 oracle with the project's libm: for each smoother (Nelder-Mead,
gradient descent, spawn search at 1/20, GETMe), the per-pass min/mean of
10 passes, the histograms after passes 1, 2 and 10, the evaluation count,
and digests of the positions after passes 1, 2 and 10; and the BATR
restructure of the same mesh (scaling 20, edges in [0.5, 1.5], 5 topology
passes, 5 refinement sweeps, glibc as the reference's libm): operation
counts, sizes and digests of the positions and tets.  Regression guards for
the oracle and, through the tests, the library.  Run from the repository
root: python tests/golden/make_c1_smoothers.py
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
    pyoracle.set_libm(1)
    out = {"workload": "Kuhn cube 20^3, jitter 0.3h, seed 1, mean ratio, metric-free, 10 passes per smoother",
           "smoothers": {}}
    for sm in ("nm", "gd", "spawn", "getme"):
        om = pyoracle.OracleMesh.from_cube(m)
        rec = {"positions_sha256": {}, "hist": {}}
        mins, means, evals = [], [], 0
        for upto, step in ((1, 1), (2, 1), (10, 8)):
            r = om.smooth(sm, step, threads=8)
            if not mins:
                mins.append(float(r["min"][0]))
                means.append(float(r["mean"][0]))
            mins += r["min"][1:].tolist()
            means += r["mean"][1:].tolist()
            evals += int(r["eval_counts"].sum())
            rec["positions_sha256"][str(upto)] = digest(om.positions())
            rec["hist"][str(upto)] = r["hist"][-1].tolist()
        rec.update({"min": mins, "mean": means, "eval_count_total": evals})
        out["smoothers"][sm] = rec
    om = pyoracle.OracleMesh.from_cube(m)
    om.set_crew("mean_ratio", "uniform", scaling=20.0)
    counts = om.restructure(5, 5, 0.5, 1.5, libm=0)
    out["restructure"] = {"scaling": 20.0, "topo_passes": 5, "refinement_sweeps": 5, "min_edge": 0.5,
                          "max_edge": 1.5, "counts": counts, "nverts": om.nverts, "ntets": om.ntet,
                          "positions_sha256": digest(om.positions()),
                          "tets_sha256": digest(om.elements()[0])}
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_smoothers.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
