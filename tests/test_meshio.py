"""Mesh files: the reference's JSON layout (Serialization/JsonSerializer.cpp:25-120,
JsonDeserializer.cpp:38-115) and the binary job files (gm_job_*).  Host only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from gpumesh_amd import GpuMeshError, cube_mesh, decode_job, encode_job, read_json, write_json

# A file in the reference's layout, written as JsonSerializer::serialize
# formats it (4-space indents, one entry per line, 6 significant digits).
REFERENCE_LAYOUT = """{
    "Model": "Cube",
    "Boundary": "Box",
    "verts": [
        [0, 0, 0],
        [1, 0, 0],
        [0, 1, 0],
        [0.333333, 0.25, 1e-07],
        [0, 0, 1]
    ],
    "topos": [
        -1,
        -1,
        -1,
        0,
        -1
    ],
    "tets": [
        [0, 1, 2, 3],
        [1, 2, 3, 4]
    ],
    "pris": [
    ],
    "hexs": [
    ]
}
"""


def test_reads_the_reference_layout(tmp_path):
    p = tmp_path / "ref.json"
    p.write_text(REFERENCE_LAYOUT)
    m = read_json(str(p))
    assert m.model == "Cube" and m.boundary == "Box"
    assert m.pos.shape == (5, 3) and m.pos[3, 0] == 0.333333 and m.pos[3, 2] == 1e-07
    assert m.fixed.tolist() == [1, 1, 1, 0, 1]
    assert m.tets.tolist() == [[0, 1, 2, 3], [1, 2, 3, 4]]
    assert m.hexs.shape == (0, 8)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_json_round_trip_is_exact(tmp_path, kind):
    m = cube_mesh(4, kind, 0.3, seed=3)
    p = tmp_path / "m.json"
    write_json(str(p), m, model="cube", boundary="Box")
    r = read_json(str(p))
    assert np.array_equal(r.pos, m.pos)                  # 17 significant digits: every bit
    assert np.array_equal(r.fixed, m.fixed.astype(np.int32))
    assert np.array_equal(r.tets, m.tets.reshape(-1, 4))
    assert np.array_equal(r.hexs, m.hexs.reshape(-1, 8))


def test_reference_digits_match_cxx_ostream(tmp_path):
    """reference_digits reproduces what `ostream << double` prints by default
    (precision 6): checked against g++'s libstdc++ on awkward values."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    vals = [0.0, -0.0, 1.0, 0.1, 1.0 / 3.0, 123456.7, 1234567.0, 1e-7, -2.5e-5, 0.30000000000000004, 1e21, 5e-324]
    src = tmp_path / "p.cpp"
    src.write_text("#include <iostream>\nint main(){double v[]={" + ",".join(repr(v) for v in vals) +
                   "};for(double x: v) std::cout << x << '\\n';}\n")
    exe = tmp_path / "p"
    subprocess.run([gxx, "-O0", str(src), "-o", str(exe)], check=True)
    cxx = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    pos = np.array(vals + [0.0] * ((-len(vals)) % 3)).reshape(-1, 3)

    class M:
        pass
    m = M()
    m.pos, m.fixed, m.tets, m.hexs = pos, np.zeros(pos.shape[0]), np.zeros((0, 4)), np.zeros((0, 8))
    p = tmp_path / "r.json"
    write_json(str(p), m, reference_digits=True)
    printed = []
    for line in p.read_text().splitlines():
        line = line.strip().rstrip(",")
        if line.startswith("[") and line.endswith("]") and "," in line:
            printed += [t.strip() for t in line[1:-1].split(",")]
    assert printed[:len(vals)] == cxx


def test_refuses_what_the_smoother_cannot_take(tmp_path):
    p = tmp_path / "bad.json"
    p.write_text(REFERENCE_LAYOUT.replace('"pris": [\n    ]', '"pris": [\n        [0, 1, 2, 3, 4, 99]\n    ]'))
    with pytest.raises(ValueError):
        read_json(str(p))                                                  # a prism vertex outside verts
    p.write_text(REFERENCE_LAYOUT.replace("        0,\n", "        3,\n"))     # a face constraint id
    with pytest.raises(ValueError):
        read_json(str(p))


def test_prism_mesh_round_trip_is_exact(tmp_path):
    from gpumesh_amd.meshgen import element_cube_mesh
    m = element_cube_mesh(3, "mixed", 0.3, 4)
    p = tmp_path / "m.json"
    write_json(str(p), m)
    r = read_json(str(p))
    assert np.array_equal(r.pos.view(np.uint64), m.pos.view(np.uint64))
    assert np.array_equal(r.fixed, m.fixed)
    assert np.array_equal(r.tets, m.tets) and np.array_equal(r.hexs, m.hexs) and np.array_equal(r.pris, m.pris)


@pytest.mark.parametrize("kind", ["tet", "hex"])
def test_job_round_trip_is_exact(kind):
    m = cube_mesh(5, kind, 0.3, seed=8)
    blob = encode_job(m, "gd", 7, "fp32")
    r, smoother, passes, precision = decode_job(blob)
    assert (smoother, passes, precision) == ("gd", 7, "fp32")
    assert np.array_equal(r.pos, m.pos) and r.pos.tobytes() == np.ascontiguousarray(m.pos).tobytes()
    assert np.array_equal(r.fixed, m.fixed.astype(np.int32))
    assert np.array_equal(r.tets, m.tets.reshape(-1, 4)) and np.array_equal(r.hexs, m.hexs.reshape(-1, 8))
    assert blob[:4] == b"GMJ1"


def test_job_damage_is_detected():
    blob = bytearray(encode_job(cube_mesh(3, "tet", 0.3, seed=1), "nm", 3))
    for where in (5, 70, len(blob) - 20):
        bad = bytearray(blob)
        bad[where] ^= 0x10
        with pytest.raises(GpuMeshError):
            decode_job(bytes(bad))
    with pytest.raises(GpuMeshError):
        decode_job(bytes(blob[:-8]))                   # truncated
    with pytest.raises(GpuMeshError):
        decode_job(b"NOPE" + bytes(blob[4:]))
