"""Python handle of the serving loop (gm_server_*): submit smoothing jobs for
meshes of any size from any thread, wait for them, read their results."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr
from .mesh import EVALUATORS, PRECISIONS, SAMPLERS, SMOOTHERS

JOB_STATES = {0: "queued", 1: "built", 2: "active", 3: "done", 4: "failed"}


class GpuServer:
    """A request loop on the current HIP device.  budget_bytes <= 0 uses the
    free device memory (less 1 GiB).  Submitted arrays are lent to the server
    (no copy) and held here until the job is built."""

    def __init__(self, budget_bytes: int = 0, max_active: int = 4096, worker_threads: int = 8,
                 evaluator: str = "mean_ratio", sampler: str = "uniform", scaling: float = 1.0, texture=None):
        """texture = (size[3], tmin[3], extent[3], texels) for sampler "texture"."""
        h = C.c_void_p()
        check(lib().gm_server_create(int(budget_bytes), int(max_active), int(worker_threads), EVALUATORS[evaluator],
                                     SAMPLERS[sampler], float(scaling), C.byref(h)), "gm_server_create")
        self._h = h
        if sampler == "texture" and texture is not None:
            size, tmin, text, texels = texture
            keep = [np.ascontiguousarray(size, np.int32), np.ascontiguousarray(tmin, np.float64),
                    np.ascontiguousarray(text, np.float64), np.ascontiguousarray(texels, np.float32)]
            check(lib().gm_server_set_texture(self._h, *[ptr(a) for a in keep]), "gm_server_set_texture")
        self._passes = {}
        self._nverts = {}
        self._lent = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().gm_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, smoother: str, patches_per_group: int, tile_elems: int) -> None:
        check(lib().gm_server_set_shape(self._h, SMOOTHERS[smoother], int(patches_per_group), int(tile_elems)),
              "gm_server_set_shape")

    def submit(self, pos, fixed, tets, hexs, smoother: str, passes: int, precision: str = "fp32_polish") -> int:
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        fixed = np.ascontiguousarray(fixed, np.int32).reshape(-1)
        tets = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
        hexs = np.ascontiguousarray(hexs, np.uint32).reshape(-1, 8)
        job = C.c_int64(0)
        check(lib().gm_server_submit(self._h, pos.shape[0], ptr(pos), ptr(fixed), tets.shape[0], ptr(tets),
                                     hexs.shape[0], ptr(hexs), SMOOTHERS[smoother], int(passes),
                                     PRECISIONS[precision], 1, C.byref(job)), "gm_server_submit")
        self._passes[job.value] = int(passes)
        self._nverts[job.value] = int(pos.shape[0])
        self._lent[job.value] = (pos, fixed, tets, hexs)
        return int(job.value)

    def submit_job(self, blob: bytes) -> int:
        """Queues a job file (gpumesh_amd.encode_job / gm_job_encode)."""
        raw = np.frombuffer(blob, np.uint8)
        counts = np.zeros(3, np.int64)
        req = np.zeros(3, np.int32)
        check(lib().gm_job_info(ptr(raw), raw.size, ptr(counts), ptr(req)), "gm_job_info")
        job = C.c_int64(0)
        check(lib().gm_server_submit_job(self._h, ptr(raw), raw.size, C.byref(job)), "gm_server_submit_job")
        self._passes[job.value] = int(req[1])
        self._nverts[job.value] = int(counts[0])
        return int(job.value)

    def submit_mesh(self, m, smoother: str, passes: int, precision: str = "fp32_polish") -> int:
        return self.submit(m.pos, m.fixed, m.tets, m.hexs, smoother, passes, precision)

    def wait(self, job: int, timeout_ms: float = -1.0) -> str:
        st = C.c_int(0)
        check(lib().gm_server_wait(self._h, int(job), float(timeout_ms), C.byref(st)), "gm_server_wait")
        if st.value != 0:
            self._lent.pop(job, None)
        return JOB_STATES[st.value]

    def result(self, job: int, release: bool = True):
        p = self._passes[job]
        pos = np.empty((self._nverts[job], 3), np.float64)
        stats = np.zeros((p + 1, 3), np.float64)
        hist = np.zeros((p + 1, 20), np.int32)
        times = np.zeros(4, np.float64)
        check(lib().gm_server_result(self._h, int(job), ptr(pos), ptr(stats), ptr(hist), ptr(times)),
              "gm_server_result")
        if release:
            self.release(job)
        return {"pos": pos, "min": stats[:, 0].copy(), "mean": stats[:, 1].copy(), "inv_sum": stats[:, 2].copy(),
                "hist": hist, "build_ms": float(times[0]), "admit_wait_ms": float(times[1]),
                "device_ms": float(times[2]), "latency_ms": float(times[3])}

    def release(self, job: int) -> None:
        check(lib().gm_server_release(self._h, int(job)), "gm_server_release")
        self._passes.pop(job, None)
        self._nverts.pop(job, None)
        self._lent.pop(job, None)

    def stats(self):
        out = np.zeros(8, np.float64)
        check(lib().gm_server_stats(self._h, ptr(out)), "gm_server_stats")
        keys = ["submitted", "done", "failed", "rounds", "peak_active", "peak_reserved_bytes", "budget_bytes",
                "round_ms"]
        return dict(zip(keys, out.tolist()))
