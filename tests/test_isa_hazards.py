"""Checks the gfx950 machine code of the built library (CPU only).

A device function (not a kernel) returns through s_setpc_b64 s[30:31], the
register pair its caller's s_swappc_b64 filled with the return address.  When
a function is long enough that the compiler expands some of its branches into
s_getpc_b64 / s_add / s_setpc_b64 sequences through s[30:31] without saving the
pair first, every return after such a branch jumps back into the function's own
body and the kernel never ends.  Kernels have no return address, so the
expansion is harmless there.  This test disassembles every gfx950 code object
of libgpumesh_hip.so with llvm-objdump and fails on a non-kernel function that
writes s[30:31] with s_getpc_b64.
"""
import os
import re
import shutil
import struct
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpumesh_amd", "libgpumesh_hip.so")
LLVM = "/opt/rocm/llvm/bin"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def code_objects(blob: bytes):
    """The device ELF images of every clang offload bundle in `blob`:
    magic, entry count, then (offset, size, triple length, triple) per entry,
    offsets relative to the bundle's start."""
    out = []
    i = 0
    while True:
        j = blob.find(MAGIC, i)
        if j < 0:
            return out
        i = j + len(MAGIC)
        (n,) = struct.unpack_from("<Q", blob, j + len(MAGIC))
        p = j + len(MAGIC) + 8
        for _ in range(n):
            off, size, tl = struct.unpack_from("<QQQ", blob, p)
            p += 24
            triple = blob[p:p + tl].decode()
            p += tl
            if "amdgcn" in triple and size:
                out.append((triple, blob[j + off:j + off + size]))


def clobbering_functions(path: str):
    """(function, count) of non-kernel functions of one code object that
    write s[30:31] with s_getpc_b64."""
    syms = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "-sW", path], capture_output=True, text=True,
                          check=True).stdout
    kernels = {f.split()[-1][:-3] for f in syms.splitlines() if f.split() and f.split()[-1].endswith(".kd")}
    dis = subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", "--no-show-raw-insn", "--no-leading-addr", path],
                         capture_output=True, text=True, check=True).stdout
    found = {}
    fn = None
    for line in dis.splitlines():
        m = re.match(r"^<(.+)>:$", line.strip()) if line.startswith("<") else None
        if m:
            fn = m.group(1)
        elif fn is not None and fn not in kernels and "s_getpc_b64 s[30:31]" in line:
            found[fn] = found.get(fn, 0) + 1
    return sorted(found.items()), len(kernels)


def test_bundle_parser_reads_a_synthetic_bundle():
    img = b"\x7fELF-device-image"
    triple = b"hipv4-amdgcn-amd-amdhsa--gfx950"
    host = b"host-x86_64-unknown-linux-gnu-"
    head = len(MAGIC) + 8 + 2 * 24 + len(triple) + len(host)
    blob = MAGIC + struct.pack("<Q", 2)
    blob += struct.pack("<QQQ", head, 0, len(host)) + host
    blob += struct.pack("<QQQ", head, len(img), len(triple)) + triple
    blob = b"junk" + blob + img
    objs = code_objects(blob)
    assert objs == [(triple.decode(), img)]


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
@pytest.mark.skipif(shutil.which(os.path.join(LLVM, "llvm-objdump")) is None, reason="llvm-objdump not present")
def test_no_function_expands_branches_through_its_return_address():
    objs = code_objects(open(LIB, "rb").read())
    assert len(objs) >= 16, "expected one code object per kernel set"
    with tempfile.TemporaryDirectory() as d:
        paths = []
        for k, (_, img) in enumerate(objs):
            p = os.path.join(d, f"co{k}.elf")
            with open(p, "wb") as f:
                f.write(img)
            paths.append(p)
        with ThreadPoolExecutor(min(8, os.cpu_count() or 1)) as ex:
            results = list(ex.map(clobbering_functions, paths))
    assert sum(n for _, n in results) > 100, "no kernels found in the code objects"
    bad = [fn for found, _ in results for fn, _ in found]
    assert not bad, f"functions that overwrite their return address in a long branch: {bad}"
