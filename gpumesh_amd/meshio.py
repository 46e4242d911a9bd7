"""Mesh files: the reference's JSON layout and the library's binary job files.

JSON (Serialization/JsonSerializer.cpp:25-120, JsonDeserializer.cpp:38-115):

    {
        "Model": "...",
        "Boundary": "...",
        "verts": [[x, y, z], ...],
        "topos": [id, ...],
        "tets": [[v0, v1, v2, v3], ...],
        "pris": [[...6...], ...],
        "hexs": [[...8...], ...]
    }

`topos` holds each vertex's boundary-constraint id (AbstractConstraint.h:65-73):
negative = fixed (MeshTopo(fixedPosition) makes VertexConstraint(-1),
Mesh.cpp:144-147), 0 = free (the volume).  The reference prints positions with
the stream's default 6 significant digits, and reading maps a fixed id back
to a boundary constraint rather than to the stored position
(AbstractBoundary.cpp:22-30), so its round trip is lossy.  `write_json` prints
17 significant digits by default (exact for doubles) and
`reference_digits=True` reproduces the reference's 6; `read_json` keeps the
stored position of a fixed vertex.  Faces and edges of a sliding boundary
(positive ids) are refused here (the library takes them through
gm_mesh_create_desc with explicit planes and segments).  Prisms are read and
written.

Job files (`encode_job` / `decode_job`, the C ABI's gm_job_*) carry a
tetrahedron/hexahedron mesh and its smoothing request bit for bit, for queues
and the serving loop.
"""
from __future__ import annotations

import json
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib, ptr
from .mesh import PRECISIONS, SMOOTHERS

_SMOOTHER_NAMES = {0: "nm", 1: "gd", 2: "spawn", 3: "getme"}
_PRECISION_NAMES = {v: k for k, v in PRECISIONS.items()}


@dataclass
class MeshData:
    pos: np.ndarray          # (V, 3) float64
    fixed: np.ndarray        # (V,) int32, 1 = fixed
    tets: np.ndarray         # (T, 4) uint32
    hexs: np.ndarray         # (H, 8) uint32
    model: str = ""
    boundary: str = ""
    pris: np.ndarray = None  # (P, 6) uint32

    @property
    def nverts(self) -> int:
        return int(self.pos.shape[0])


def _num(x: float, digits: int) -> str:
    """A double as C++'s ostream prints it with precision `digits` (the
    %g conversion: shortest of fixed and scientific, trailing zeros dropped)."""
    return format(float(x), f".{digits}g")


def write_json(path: str, m, model: str = "", boundary: str = "", reference_digits: bool = False) -> None:
    """Writes `m` (anything with pos, fixed, tets, hexs and optionally pris)
    in the reference's JSON layout; 17 significant digits unless
    reference_digits."""
    digits = 6 if reference_digits else 17
    pos = np.asarray(m.pos, np.float64).reshape(-1, 3)
    fixed = np.asarray(m.fixed).reshape(-1)
    tets = np.asarray(m.tets, np.int64).reshape(-1, 4)
    hexs = np.asarray(m.hexs, np.int64).reshape(-1, 8)
    pris = getattr(m, "pris", None)
    pris = np.zeros((0, 6), np.int64) if pris is None else np.asarray(pris, np.int64).reshape(-1, 6)
    ind = "    "
    out = ["{\n", f'{ind}"Model": "{model}",\n', f'{ind}"Boundary": "{boundary}",\n', f'{ind}"verts": [\n']
    n = pos.shape[0]
    for i, p in enumerate(pos):
        out.append(f"{ind}{ind}[{_num(p[0], digits)}, {_num(p[1], digits)}, {_num(p[2], digits)}]"
                   + ("," if i != n - 1 else "") + "\n")
    out.append(f"{ind}],\n{ind}\"topos\": [\n")
    for i, f in enumerate(fixed):
        out.append(f"{ind}{ind}{-1 if f else 0}" + ("," if i != n - 1 else "") + "\n")
    for tag, arr in (("tets", tets), ("pris", pris), ("hexs", hexs)):
        out.append(f"{ind}],\n{ind}\"{tag}\": [\n")
        k = arr.shape[0]
        for i, row in enumerate(arr):
            out.append(f"{ind}{ind}[" + ", ".join(str(int(v)) for v in row) + "]" + ("," if i != k - 1 else "") + "\n")
    out.append(f"{ind}]\n}}\n")
    with open(path, "w") as f:
        f.write("".join(out))


def read_json(path: str) -> MeshData:
    """Reads a mesh in the reference's JSON layout."""
    with open(path) as f:
        doc = json.load(f)
    unknown = set(doc) - {"Model", "Boundary", "verts", "topos", "tets", "pris", "hexs"}
    if unknown:
        raise ValueError(f"unknown tag(s) in json mesh: {sorted(unknown)}")
    pos = np.asarray(doc.get("verts", []), np.float64).reshape(-1, 3)
    topos = np.asarray(doc.get("topos", [0] * pos.shape[0]), np.int64).reshape(-1)
    if topos.size != pos.shape[0]:
        raise ValueError("topos and verts differ in length")
    if np.any(topos > 0):
        raise ValueError("sliding boundary constraints (positive topo ids) are not supported: only fixed "
                         "(negative) and free (0) vertices")
    tets = np.asarray(doc.get("tets", []), np.int64).reshape(-1, 4)
    pris = np.asarray(doc.get("pris", []), np.int64).reshape(-1, 6)
    hexs = np.asarray(doc.get("hexs", []), np.int64).reshape(-1, 8)
    for name, arr in (("tets", tets), ("pris", pris), ("hexs", hexs)):
        if arr.size and (arr.min() < 0 or arr.max() >= pos.shape[0]):
            raise ValueError(f"{name} reference a vertex outside verts")
    return MeshData(pos, (topos < 0).astype(np.int32), tets.astype(np.uint32), hexs.astype(np.uint32),
                    str(doc.get("Model", "")), str(doc.get("Boundary", "")), pris.astype(np.uint32))


def encode_job(m, smoother: str, passes: int, precision: str = "fp32_polish") -> bytes:
    """A job file (gm_job_encode) for mesh `m` and the smoothing request."""
    pos = np.ascontiguousarray(m.pos, np.float64).reshape(-1, 3)
    fixed = np.ascontiguousarray(m.fixed, np.int32).reshape(-1)
    tets = np.ascontiguousarray(m.tets, np.uint32).reshape(-1, 4)
    hexs = np.ascontiguousarray(m.hexs, np.uint32).reshape(-1, 8)
    L = lib()
    n = L.gm_job_size(pos.shape[0], tets.shape[0], hexs.shape[0])
    if n < 0:
        raise ValueError("bad element counts")
    buf = np.zeros(n, np.uint8)
    check(L.gm_job_encode(pos.shape[0], ptr(pos), ptr(fixed), tets.shape[0], ptr(tets), hexs.shape[0], ptr(hexs),
                          SMOOTHERS[smoother], int(passes), PRECISIONS[precision], ptr(buf), n), "gm_job_encode")
    return buf.tobytes()


def decode_job(blob: bytes):
    """(MeshData, smoother, passes, precision) of a job file."""
    raw = np.frombuffer(blob, np.uint8)
    counts = np.zeros(3, np.int64)
    req = np.zeros(3, np.int32)
    L = lib()
    check(L.gm_job_info(ptr(raw), raw.size, ptr(counts), ptr(req)), "gm_job_info")
    nv, nt, nh = (int(x) for x in counts)
    pos = np.zeros((nv, 3), np.float64)
    fixed = np.zeros(nv, np.int32)
    tets = np.zeros((nt, 4), np.uint32)
    hexs = np.zeros((nh, 8), np.uint32)
    check(L.gm_job_decode(ptr(raw), raw.size, ptr(pos), ptr(fixed), ptr(tets), ptr(hexs)), "gm_job_decode")
    return (MeshData(pos, fixed, tets, hexs), _SMOOTHER_NAMES[int(req[0])], int(req[1]),
            _PRECISION_NAMES[int(req[2])])
