"""Metric fields for the anisotropic workloads.

shell_metric(p, K): the spherical-shell field of the anisotropic hex
workload, M(x) = K^2 (a^2 I + (1 - a^2) d d^T / |d|^2) with d = x - c,
c = (0.5, 0.5, 0.5), a = 1 + 9 exp(-50 (|d| - 0.3)^2): unit metric length
along the radius, a^2 across it (a 10:1 shell at |d| = 0.3), scaled by the
mesh size K = 1/h.  The device evaluates the same formula (gm_quality.h,
shellMetric).

shell_texture(pos, n, K): the same field sampled at the cell centres of an
n^3 grid over the mesh bounding box and stored as FP32 symmetric texels
(m00 m01 m02 m11 m12 m22), the layout the texture sampler reads; cell centre
placement and clamping follow TextureSampler (Samplers/TextureSampler.cpp:275-448,
450-488), with the grid filled directly from the analytic field.
"""
from __future__ import annotations

import numpy as np


def shell_metric(p: np.ndarray, K: float) -> np.ndarray:
    p = np.asarray(p, np.float64)
    d = p - 0.5
    r2 = np.sum(d * d, axis=-1)
    r = np.sqrt(r2)
    a = 1.0 + 9.0 * np.exp(-50.0 * (r - 0.3) ** 2)
    a2 = a * a
    with np.errstate(divide="ignore", invalid="ignore"):
        w = np.where(r2 > 0, (1.0 - a2) / np.where(r2 > 0, r2, 1.0), 0.0)
    k2 = K * K
    m = np.empty(p.shape[:-1] + (6,), np.float64)
    m[..., 0] = k2 * (a2 + w * d[..., 0] * d[..., 0])
    m[..., 1] = k2 * (w * d[..., 0] * d[..., 1])
    m[..., 2] = k2 * (w * d[..., 0] * d[..., 2])
    m[..., 3] = k2 * (a2 + w * d[..., 1] * d[..., 1])
    m[..., 4] = k2 * (w * d[..., 1] * d[..., 2])
    m[..., 5] = k2 * (a2 + w * d[..., 2] * d[..., 2])
    return m


def shell_texture(pos: np.ndarray, n: int, K: float):
    """Returns (size[3] int32, tmin[3], text[3], texels float32 (n^3 * 6))."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    tmin = pos.min(axis=0)
    text = pos.max(axis=0) - tmin
    size = np.array([n, n, n], np.int32)
    cs = text / size
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    centres = tmin + cs * (np.stack([i, j, k], axis=-1) + 0.5)
    texels = shell_metric(centres, K).astype(np.float32)   # (k, j, i, 6)
    return size, tmin, text, np.ascontiguousarray(texels.reshape(-1))
