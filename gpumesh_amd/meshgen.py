"""Synthetic structured meshes of the unit cube with seeded jitter.

The smoothing workloads are quoted on cubes of n^3 cells:
  * Kuhn tetrahedral cubes: every cell split into the 6 tetrahedra that share
    its (0,0,0)-(1,1,1) diagonal;
  * hexahedral cubes: one hexahedron per cell.

Vertex (i, j, k) sits at (i/n, j/n, k/n) and is numbered x-fastest,
i + (n+1)(j + (n+1)k), the numbering of the reference's grid mesher
(Meshers/DebugMesher.cpp:173-181).  Hexahedra list v0..v3 counter-clockwise
on the bottom face and v4..v7 on the top face (DebugMesher.cpp:285-293).
Tetrahedra are oriented so that det(v3-v0, v3-v1, v3-v2) > 0, the sign
convention under which MeanRatioEvaluator::tetQuality
(Evaluators/MeanRatioEvaluator.cpp:34-50) returns a positive quality.

Boundary vertices are fixed (a fixed-position constraint, Mesh.cpp:144-147);
interior vertices get a per-axis jitter drawn uniformly from [-a h, +a h]
with a portable counter-based generator (SplitMix64 of seed and counter,
53-bit mantissa), so every platform generates identical coordinates.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser applied element-wise to uint64 counters."""
    with np.errstate(over="ignore"):
        z = z.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, counters: np.ndarray) -> np.ndarray:
    """U[0, 1) doubles from (seed, counter): 53 high bits of SplitMix64."""
    with np.errstate(over="ignore"):
        key = np.uint64(seed) * np.uint64(0xD1B54A32D192ED03) + counters.astype(np.uint64)
    bits = splitmix64(key) >> np.uint64(11)
    return bits.astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class CubeMesh:
    """A generated mesh: positions (V,3) float64, fixed flags (V,) int32,
    tets (T,4) uint32, hexs (H,8) uint32, prisms (P,6) uint32 and its grid
    parameters."""
    pos: np.ndarray
    fixed: np.ndarray
    tets: np.ndarray
    hexs: np.ndarray
    n: int
    kind: str
    seed: int
    jitter: float
    pris: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.uint32))

    @property
    def nverts(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nelems(self) -> int:
        return int(self.tets.shape[0] + self.pris.shape[0] + self.hexs.shape[0])

    @property
    def nmovable(self) -> int:
        return int(self.nverts - int(self.fixed.sum()))


# The 6 Kuhn tetrahedra of a cell, as cell-corner codes bx + 2 by + 4 bz:
# monotone lattice paths 0 -> e_a -> e_a + e_b -> 7.
_KUHN_PATHS = [(0, 1, 3, 7), (0, 1, 5, 7), (0, 2, 3, 7), (0, 2, 6, 7), (0, 4, 5, 7), (0, 4, 6, 7)]


def _corner_offsets(n: int) -> np.ndarray:
    w = n + 1
    return np.array([(c & 1) + ((c >> 1) & 1) * w + ((c >> 2) & 1) * w * w for c in range(8)], dtype=np.int64)


def _kuhn_local() -> np.ndarray:
    """Per-cell tetrahedra with positive orientation (swap v1, v2 if needed)."""
    corner_xyz = np.array([[(c & 1), (c >> 1) & 1, (c >> 2) & 1] for c in range(8)], dtype=np.float64)
    out = []
    for path in _KUHN_PATHS:
        v = list(path)
        p = corner_xyz[v]
        det = np.linalg.det(np.stack([p[3] - p[0], p[3] - p[1], p[3] - p[2]], axis=1))
        if det < 0:
            v[1], v[2] = v[2], v[1]
        out.append(v)
    return np.array(out, dtype=np.int64)


def _grid(n: int):
    w = n + 1
    k, j, i = np.meshgrid(np.arange(w), np.arange(w), np.arange(w), indexing="ij")
    i = i.reshape(-1)
    j = j.reshape(-1)
    k = k.reshape(-1)
    pos = np.stack([i / n, j / n, k / n], axis=1).astype(np.float64)
    boundary = (i == 0) | (i == n) | (j == 0) | (j == n) | (k == 0) | (k == n)
    return pos, boundary


def _cell_base(n: int) -> np.ndarray:
    w = n + 1
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return (ii + jj * w + kk * w * w).reshape(-1).astype(np.int64)


def _dot3(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Row-wise dot product summed left to right, (a0 b0 + a1 b1) + a2 b2.
    (np.einsum may fuse or reorder the sum depending on the host CPU, which
    would make generated meshes machine-dependent.)"""
    return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]


def _cross3(b: np.ndarray, c: np.ndarray) -> np.ndarray:
    return np.stack([b[:, 1] * c[:, 2] - b[:, 2] * c[:, 1], b[:, 2] * c[:, 0] - b[:, 0] * c[:, 2],
                     b[:, 0] * c[:, 1] - b[:, 1] * c[:, 0]], axis=1)


def _tet_dets(pos: np.ndarray, tets: np.ndarray) -> np.ndarray:
    p = pos[tets]
    a = p[:, 3] - p[:, 0]
    b = p[:, 3] - p[:, 1]
    c = p[:, 3] - p[:, 2]
    return _dot3(a, _cross3(b, c))


def _hex_min_corner_det(pos: np.ndarray, hexs: np.ndarray) -> np.ndarray:
    """Minimum over the 8 corner Jacobian determinants, using the corner
    frames of MeanRatioEvaluator::hexQuality (MeanRatioEvaluator.cpp:127-134)."""
    p = pos[hexs]

    def e(i, j):
        return p[:, j] - p[:, i]

    frames = [(e(0, 1), e(0, 4), -e(0, 3)), (e(0, 1), e(1, 2), e(1, 5)),
              (e(1, 2), e(2, 6), -e(2, 3)), (e(0, 3), e(2, 3), e(3, 7)),
              (e(0, 4), e(4, 5), e(4, 7)), (e(1, 5), -e(5, 6), e(4, 5)),
              (e(2, 6), e(5, 6), e(6, 7)), (e(3, 7), e(6, 7), -e(4, 7))]
    dets = [_dot3(a, _cross3(b, c)) for a, b, c in frames]
    return np.min(np.stack(dets, axis=1), axis=1)


# The two prisms of a cell: the bottom square split along its (0,0)-(1,1)
# diagonal into counter-clockwise triangles (viewed from +z), extruded to the
# top square; cell-corner codes bx + 2 by + 4 bz.  Vertices 0-2 form the
# bottom triangle and 3-5 the top one above them (MeshPri's order).
_PRISM_CELL = [(0, 1, 3, 4, 5, 7), (0, 3, 2, 4, 7, 6)]


def _pri_min_corner_det(pos: np.ndarray, pris: np.ndarray) -> np.ndarray:
    """Minimum over the 6 corner frames of MeanRatioEvaluator::priQuality
    (Evaluators/MeanRatioEvaluator.cpp:70-75) before the right factor
    Fr_PRI_INV, whose determinant is positive."""
    p = pos[pris]

    def e(i, j):
        return p[:, j] - p[:, i]

    frames = [(-e(0, 1), e(2, 0), e(0, 3)), (-e(1, 2), e(0, 1), e(1, 4)), (-e(2, 0), e(1, 2), e(2, 5)),
              (-e(3, 4), e(5, 3), e(0, 3)), (-e(4, 5), e(3, 4), e(1, 4)), (-e(5, 3), e(4, 5), e(2, 5))]
    dets = [_dot3(a, _cross3(b, c)) for a, b, c in frames]
    return np.min(np.stack(dets, axis=1), axis=1)


def element_cube_mesh(n: int, layout: str = "prism", jitter: float = 0.3, seed: int = 1,
                      min_det_fraction: float = 0.05) -> CubeMesh:
    """An n^3-cell unit-cube mesh with prisms: layout "prism" splits every
    cell into two prisms; layout "mixed" cuts the cube into three slabs
    along x, hexahedra for i < n/3, prisms for n/3 <= i < 2n/3 and Kuhn
    tetrahedra beyond (the slabs meet without matching faces, which the
    smoothers do not need).  Jitter and its repair as cube_mesh_numpy."""
    if n < 1:
        raise ValueError("n must be >= 1")
    if layout not in ("prism", "mixed"):
        raise ValueError("layout must be 'prism' or 'mixed'")
    h = 1.0 / n
    base, boundary = _grid(n)
    V = base.shape[0]
    vid = np.arange(V, dtype=np.uint64)
    u = np.stack([uniform01(seed, vid * np.uint64(3) + np.uint64(d)) for d in range(3)], axis=1)
    disp = (2.0 * u - 1.0) * (jitter * h)
    disp[boundary] = 0.0
    cells = _cell_base(n)
    ci = np.arange(n ** 3) % n
    off = _corner_offsets(n)
    if layout == "prism":
        hex_cells = np.zeros(0, np.int64)
        pri_cells = cells
        tet_cells = np.zeros(0, np.int64)
    else:
        hex_cells = cells[ci < n // 3]
        pri_cells = cells[(ci >= n // 3) & (ci < (2 * n) // 3)]
        tet_cells = cells[ci >= (2 * n) // 3]
    tets = (tet_cells[:, None, None] + off[_kuhn_local()][None, :, :]).reshape(-1, 4)
    pris = (pri_cells[:, None, None] + off[np.array(_PRISM_CELL)][None, :, :]).reshape(-1, 6)
    hexs = hex_cells[:, None] + off[[0, 1, 3, 2, 4, 5, 7, 6]][None, :]
    ref_tet = h ** 3
    checks = [(tets, _tet_dets, ref_tet), (pris, _pri_min_corner_det, h ** 3), (hexs, _hex_min_corner_det, h ** 3)]
    scale = np.ones(V)
    for _ in range(64):
        pos = base + disp * scale[:, None]
        bad_verts = [np.unique(el[fn(pos, el) < min_det_fraction * ref].reshape(-1)) for el, fn, ref in checks
                     if el.shape[0]]
        bad = np.unique(np.concatenate(bad_verts)) if bad_verts else np.zeros(0, np.int64)
        if bad.size == 0:
            break
        scale[bad] *= 0.5
    else:  # pragma: no cover - scale -> 0 gives the regular grid
        raise RuntimeError("jitter repair did not converge")
    pos = base + disp * scale[:, None]
    return CubeMesh(pos=np.ascontiguousarray(pos), fixed=boundary.astype(np.int32),
                    tets=np.ascontiguousarray(tets.astype(np.uint32).reshape(-1, 4)),
                    hexs=np.ascontiguousarray(hexs.astype(np.uint32).reshape(-1, 8)), n=n, kind=layout,
                    seed=seed, jitter=jitter, pris=np.ascontiguousarray(pris.astype(np.uint32).reshape(-1, 6)))


def cube_mesh(n: int, kind: str = "tet", jitter: float = 0.3, seed: int = 1,
              min_det_fraction: float = 0.05) -> CubeMesh:
    """Builds an n^3-cell unit-cube mesh ("tet" Kuhn split or "hex") with the
    library's native generator (gm_cube_mesh); the output is identical to
    cube_mesh_numpy, the restatement below."""
    if n < 1:
        raise ValueError("n must be >= 1")
    if kind not in ("tet", "hex"):
        raise ValueError("kind must be 'tet' or 'hex'")
    from ._lib import check, lib, ptr
    w = n + 1
    pos = np.empty((w ** 3, 3), np.float64)
    fixed = np.empty(w ** 3, np.int32)
    per = 4 if kind == "tet" else 8
    ne = 6 * n ** 3 if kind == "tet" else n ** 3
    elems = np.empty((ne, per), np.uint32)
    check(lib().gm_cube_mesh(int(n), 0 if kind == "tet" else 1, float(jitter), int(seed) & 0xFFFFFFFFFFFFFFFF,
                             float(min_det_fraction), ptr(pos), ptr(fixed), ptr(elems)), "gm_cube_mesh")
    empty4 = np.zeros((0, 4), np.uint32)
    empty8 = np.zeros((0, 8), np.uint32)
    return CubeMesh(pos=pos, fixed=fixed, tets=elems if kind == "tet" else empty4,
                    hexs=elems if kind == "hex" else empty8, n=n, kind=kind, seed=seed, jitter=jitter)


def cube_mesh_numpy(n: int, kind: str = "tet", jitter: float = 0.3, seed: int = 1,
                    min_det_fraction: float = 0.05) -> CubeMesh:
    """Builds an n^3-cell unit-cube mesh ("tet" Kuhn split or "hex").

    Interior vertices are displaced by U[-jitter h, +jitter h] per axis.
    Elements whose orientation determinant falls below `min_det_fraction`
    of the unjittered value are repaired by halving the jitter of their
    interior vertices, repeated until none remain (the no-inversion clamp).
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    if kind not in ("tet", "hex"):
        raise ValueError("kind must be 'tet' or 'hex'")
    h = 1.0 / n
    base, boundary = _grid(n)
    V = base.shape[0]
    vid = np.arange(V, dtype=np.uint64)
    u = np.stack([uniform01(seed, vid * np.uint64(3) + np.uint64(d)) for d in range(3)], axis=1)
    disp = (2.0 * u - 1.0) * (jitter * h)
    disp[boundary] = 0.0

    cells = _cell_base(n)
    off = _corner_offsets(n)
    if kind == "tet":
        local = _kuhn_local()
        elems = (cells[:, None, None] + off[local][None, :, :]).reshape(-1, 4)
        ref = _tet_dets(base, elems[:6])
        ref_det = float(np.min(np.abs(ref)))
        det_fn = _tet_dets
    else:
        hex_codes = [0, 1, 3, 2, 4, 5, 7, 6]
        elems = cells[:, None] + off[hex_codes][None, :]
        ref_det = h ** 3
        det_fn = _hex_min_corner_det

    scale = np.ones(V)
    for _ in range(64):
        pos = base + disp * scale[:, None]
        bad = det_fn(pos, elems) < min_det_fraction * ref_det
        if not bad.any():
            break
        verts = np.unique(elems[bad].reshape(-1))
        scale[verts] *= 0.5
    else:  # pragma: no cover - the loop always converges (scale -> 0 gives the regular grid)
        raise RuntimeError("jitter repair did not converge")

    pos = base + disp * scale[:, None]
    fixed = boundary.astype(np.int32)
    tets = elems.astype(np.uint32) if kind == "tet" else np.zeros((0, 4), np.uint32)
    hexs = elems.astype(np.uint32) if kind == "hex" else np.zeros((0, 8), np.uint32)
    return CubeMesh(pos=np.ascontiguousarray(pos), fixed=fixed, tets=np.ascontiguousarray(tets),
                    hexs=np.ascontiguousarray(hexs), n=n, kind=kind, seed=seed, jitter=jitter)


def box_boundary(m: CubeMesh, slide_jitter: float = 0.2, seed: int = 7):
    """The cube mesh on the reference's BoxBoundary [-1, 1]^3
    (Boundaries/BoxBoundary.cpp:7-36): corner vertices fixed, edge vertices
    sliding on the box's 12 edges (segments), face vertices on its 6 faces
    (planes), interior vertices free.  Edge and face vertices get a seeded
    jitter of +-slide_jitter h along their constraint.  Returns (pos, fixed,
    constraint, constraints) with constraints as rows (kind, a, b): kind 1 a
    plane through a with normal b, kind 2 the segment from a to b."""
    n = m.n
    pos = m.pos * 2.0 - 1.0
    h = 2.0 / n
    V = pos.shape[0]
    on = np.isclose(np.abs(pos), 1.0)                 # (V, 3): coordinate on a box face
    count = on.sum(axis=1)
    constraints = []
    index = {}
    # Faces: xNeg, xPos, yNeg, yPos, zNeg, zPos (BoxBoundary.cpp:29-34).
    for axis in range(3):
        for sgn in (-1.0, 1.0):
            p = [0.0, 0.0, 0.0]
            p[axis] = sgn
            index[("face", axis, sgn)] = len(constraints)
            constraints.append((1, tuple(p), tuple(p)))
    # Edges: along axis `axis`, at the other two coordinates' signs.
    for axis in range(3):
        o1, o2 = [a for a in range(3) if a != axis]
        for s1 in (-1.0, 1.0):
            for s2 in (-1.0, 1.0):
                a = [0.0, 0.0, 0.0]
                b = [0.0, 0.0, 0.0]
                a[axis], b[axis] = -1.0, 1.0
                a[o1] = b[o1] = s1
                a[o2] = b[o2] = s2
                index[("edge", axis, s1, s2)] = len(constraints)
                constraints.append((2, tuple(a), tuple(b)))
    fixed = (count == 3).astype(np.int32)
    constraint = -np.ones(V, np.int32)
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-slide_jitter * h, slide_jitter * h, size=(V, 3))
    for v in np.where(count == 1)[0]:
        axis = int(np.where(on[v])[0][0])
        constraint[v] = index[("face", axis, float(np.sign(pos[v, axis])))]
        j = jit[v].copy()
        j[axis] = 0.0
        pos[v] += j
    for v in np.where(count == 2)[0]:
        free_axis = int(np.where(~on[v])[0][0])
        o1, o2 = [a for a in range(3) if a != free_axis]
        constraint[v] = index[("edge", free_axis, float(np.sign(pos[v, o1])), float(np.sign(pos[v, o2])))]
        pos[v, free_axis] += jit[v, free_axis]
    return pos, fixed, constraint, constraints
